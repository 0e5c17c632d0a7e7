/*
 * or_oracle.h -- CPU restatement of the OpenRadioss Engine explicit hot path
 * (8-node 1-point hex SFORC3 chain, law-2 Johnson-Cook, Parith/ON skyline
 * assembly, ACCELE / BCS10 / FIXVEL / VELOCITY / DEPLA, element time step).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The product
 * (openradioss_amd/) never links, loads or calls it.
 *
 * Every routine follows the operation order of the reference routine it
 * cites, in IEEE double precision with floating-point contraction disabled
 * (the reference release flags are -ffp-contract=off -frounding-math,
 * engine/CMake_Compilers/cmake_linux64_gf.txt:163).
 */
#ifndef OR_ORACLE_H
#define OR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Elementary functions of the hardening laws and the bulk viscosity.
 * Default build: the host C library, as the reference engine calls it (X**Y,
 * LOG, EXP of the Fortran sources).  Built with -DOR_CR_MATH: correctly
 * rounded pow / log / exp (or_crmath.cpp), i.e. the reference algorithm under
 * a correctly rounded libm.  The two builds differ only where the host
 * library's result is not the correctly rounded one (about 1 call in 1000 for
 * glibc 2.35 exp and pow). */
#ifdef OR_CR_MATH
double or_cr_pow(double x, double y);
double or_cr_log(double x);
double or_cr_exp(double x);
#define OR_POW or_cr_pow
#define OR_LOG or_cr_log
#define OR_EXP or_cr_exp
#else
#define OR_POW pow
#define OR_LOG log
#define OR_EXP exp
#endif

/* Call sites of the elementary functions.  The counting build (-DOR_COUNT_MATH) returns
 * the host library's value, so it follows the reference's trajectory, and counts per site
 * the calls whose correctly rounded value differs from it (or_math_site_counts). */
enum {
  OR_SITE_MQVISCB_AL = 0, /* mqviscb.F: AL = VOL**THIRD */
  OR_SITE_SHVIS3_FCL,     /* shvis3.F: VOL**TWO_THIRD */
  OR_SITE_M2LAW_AK,       /* m2law.F: EPXE**CN (yield stress) */
  OR_SITE_M2LAW_QH,       /* m2law.F: EPXE**(CN-1) or EPXE**(1-CN) (hardening modulus) */
  OR_SITE_M2LAW_RATE,     /* m2law.F: LOG(EPSD/EPDR), TSTAR**MT */
  OR_SITE_M2CPLR,         /* m2cplr.F: EXP(CN*LOG(PLA)), LOG of the rate ratio */
  OR_NSITE
};
#ifdef OR_COUNT_MATH
double or_count_pow(int site, double x, double y);
double or_count_log(int site, double x);
double or_count_exp(int site, double x);
void or_math_site_counts(int64_t *out); /* [OR_NSITE][2]: calls, calls whose results differ */
void or_math_site_reset(void);
#define OR_POWS(s, x, y) or_count_pow(s, x, y)
#define OR_LOGS(s, x) or_count_log(s, x)
#define OR_EXPS(s, x) or_count_exp(s, x)
#else
#define OR_POWS(s, x, y) OR_POW(x, y)
#define OR_LOGS(s, x) OR_LOG(x)
#define OR_EXPS(s, x) OR_EXP(x)
#endif

/* Material (law 2, Johnson-Cook form) + solid property + run control. */
typedef struct {
  double rho0;     /* mat_param%rho        (hm_read_mat02_jc.F90: pm(1)=rhor) */
  double bulk;     /* mat_param%bulk  = E/(3(1-2nu)) */
  double shear;    /* mat_param%shear = E/(2(1+nu))  */
  double ca, cb, cn;   /* uparam(1..3) */
  double epmx;     /* uparam(4): failure plastic strain (default EP20) */
  double sigmx;    /* uparam(5): max stress (default EP20) */
  double cc;       /* uparam(6): strain-rate coefficient */
  double epdr;     /* uparam(7): reference strain rate */
  double fisokin;  /* uparam(8): must be 0 (isotropic hardening) */
  double asrate;   /* uparam(9): 2*pi*fcut or EP20 */
  double m_exp;    /* uparam(10): thermal exponent */
  int    icc, vp, israte, iform;
  /* /PROP/SOLID Isolid=1 (hm_read_prop14.F:355-369) */
  double qa, qb;   /* GEO(14), GEO(15) */
  double qh;       /* GEO(13) hourglass coefficient */
  double vns1, vns2; /* GEO(16), GEO(17) */
  double dtfac;    /* DTFAC1(1) */
  double dtmx;     /* DTMX (freform.F:589 default EP20) */
} or_hex_params;

/* Scalars carried from cycle to cycle (resol.F:2721-2722, 6352, 6494-6497, 8599-8613). */
typedef struct {
  double tt, dt1, dt2, dt12, dt2old;
  int64_t ncycle;
  int    mstop;    /* set when an element reports a non-positive Jacobian */
} or_time_state;

/* SFORC3 chain for elements [e0,e1), writing skyline rows.  Returns the
 * element time-step minimum folded into *dt2t. */
/* or_hex_forces that also writes each element's time-step candidate for NELTS (dtel, may be NULL) */
void or_hex_forces_dt(int64_t e0, int64_t e1, int64_t ne, const int32_t *conn, const double *x, const double *v,
                      double *sig, double *eint, double *rho, const double *vol0, double *pla, double *epsd,
                      double *offg, double *qvis, double *ehou, const int64_t *iads, double *fsky,
                      const or_hex_params *p, double dt1, double *dt2t, int *mstop, double *dtel);
void or_hex_forces(int64_t e0, int64_t e1, int64_t ne,
                   const int32_t *conn,            /* [ne*8] 0-based node ids */
                   const double *x, const double *v, /* [3*nn] AoS X(3,NUMNOD) */
                   double *sig,                    /* [6*ne] component-major */
                   double *eint, double *rho, const double *vol0,
                   double *pla, double *epsd, double *offg, double *qvis,
                   double *ehou,                   /* [ne] accumulated hourglass energy or NULL */
                   const int64_t *iads,            /* [ne*8] 0-based skyline rows */
                   double *fsky,                   /* [8*lsky] FSKY(8,LSKY) */
                   const or_hex_params *p, double dt1,
                   double *dt2t, int *mstop);

/* ASSPAR4 (asspar4.F:164-180) with A and STIFN starting from zero. */
void or_asspar4(int64_t nn, const int64_t *adsky, const double *fsky,
                double *a, double *ar, double *stifn, double *stifr);

/* One full cycle of the hex path.  bcs[n]: BCS code (bcs1.F LCOD bits 4=x,2=y,1=z)
 * nimp impvel entries (node, dof 0..2, value). */
void or_cycle_hex(int64_t nn, int64_t ne, int64_t lsky,
                  const int32_t *conn, double *x, double *v, double *a, double *d,
                  const double *ms, double *stifn,
                  double *sig, double *eint, double *rho, const double *vol0,
                  double *pla, double *epsd, double *offg, double *qvis, double *ehou,
                  const int64_t *iads, const int64_t *adsky, double *fsky,
                  const int32_t *bcs, int64_t nimp, const int32_t *imp_node,
                  const int32_t *imp_dof, const double *imp_val,
                  const or_hex_params *p, or_time_state *ts, int nthreads);

/* Energies: internal sum(EINT*VOL0), kinetic 0.5*sum(MS*|V|^2). */
void or_energy(int64_t nn, int64_t ne, const double *eint, const double *vol0,
               const double *ms, const double *v, double out[2]);

/* Exposed leaves for unit checks against the compiled reference. */
double or_const(int which);

/* ---------------------------------------------------------------- Q4 shell */
/* Law 2 as SIGEPS02C reads it, the PM words CCOEF3 reads and /PROP/SHELL
 * (hm_read_mat02_jc.F90:236-290, hm_read_prop01.F:156-245). */
typedef struct {
  double rho;      /* PM(1) */
  double young;    /* PM(20), mat_param%young */
  double nu;       /* PM(21) */
  double shear;    /* PM(22) */
  double ssp;      /* PM(27) = SQRT(E/RHO) */
  double ca, cb, cn, epmx, ymax, cc, epdr, m_exp; /* uparam(1..7,10) */
  double asrate;   /* PM(9): 2*pi*fcut (strain-rate filter) */
  int    icc, vp, israte, iform;
  int    npt;      /* through-thickness points, 2..10 */
  int    ipla;     /* Iplas: 0 radial return, 1 Newton (3 iterations), 2 normal projection */
  int    ithk;     /* Ithick: 1 = thickness change enters the element */
  int    ishear;   /* GEO(37) */
  double hm, hf, hr; /* GEO(13..15) hourglass coefficients */
  double fsh;      /* GEO(38) shear factor (5/6) */
  double dtfac;    /* DTFAC1(3) */
  double dtmin;    /* DTMIN1(3) */
} or_q4_params;

/* GBUF / LBUF state of ne Q4 elements, component-major like ELBUF (SIG(NEL,5) ...). */
typedef struct {
  int64_t ne;
  double *offg, *thk;      /* GBUF%OFF, GBUF%THK */
  const double *thke;      /* initial thickness (THKE) */
  double *forc;            /* GBUF%FOR  [5][ne] */
  double *mom;             /* GBUF%MOM  [3][ne] */
  double *gstr;            /* GBUF%STRA [8][ne] */
  double *hourg;           /* GBUF%HOURG [5][ne] */
  double *eint;            /* GBUF%EINT [2][ne] (membrane, bending) */
  double *epsd;            /* GBUF%EPSD [ne] */
  double *sig;             /* LBUF%SIG per point [npt][5][ne] */
  double *pla;             /* LBUF%PLA [npt][ne] */
  double *epsd_ip;         /* LBUF%EPSD [npt][ne] */
} or_q4_state;

/* or_q4_forces that also writes each element's time-step candidate for NELTS (dtel, may be NULL) */
void or_q4_forces_dt(int64_t e0, int64_t e1, const int32_t *conn, const double *x, const double *v,
                     const double *vr, or_q4_state *s, const int64_t *iadc, double *fsky,
                     const or_q4_params *p, double dt1, double *dt2t, double *dtel);
/* ... and each element's hourglass energy of the cycle, CHVIS3's EHOU (ehou, may be NULL) */
void or_q4_forces_dte(int64_t e0, int64_t e1, const int32_t *conn, const double *x, const double *v,
                      const double *vr, or_q4_state *s, const int64_t *iadc, double *fsky,
                      const or_q4_params *p, double dt1, double *dt2t, double *dtel, double *ehou);
void or_q4_forces(int64_t e0, int64_t e1, const int32_t *conn /* [ne*4] */, const double *x, const double *v,
                  const double *vr, or_q4_state *s, const int64_t *iadc /* [ne*4] */, double *fsky,
                  const or_q4_params *p, double dt1, double *dt2t);
double or_shell_const(int which);
double or_shell_table(int which, int ipt, int npt);

/* A mixed hex + Q4 model with 6-DOF nodes.  code[n] bits: BCS translations
 * 4=x 2=y 1=z, BCS rotations 32=x 16=y 8=z, imposed velocity 256=x 128=y 64=z. */
typedef struct {
  int64_t nn;
  double *x, *v, *a, *d, *vr, *ar, *stifn, *stifr;
  const double *ms, *in;   /* in may be NULL (no rotational DOF) */
  const int32_t *code;
  const double *vimp;      /* [3*nn] imposed velocities */
  const int64_t *adsky;
  double *fsky;            /* FSKY(8,LSKY) */
  double dtmx;
  int64_t ne_hex;
  const int32_t *hconn; const int64_t *hiads;
  double *hsig, *heint, *hrho, *hpla, *hepsd, *hoffg, *hqvis;
  const double *hvol0;
  const or_hex_params *hp;
  int64_t ne_q4;
  const int32_t *qconn; const int64_t *qiadc;
  or_q4_state q;
  const or_q4_params *qp;
  /* if non-NULL, or_cycle_model writes the energies of ECRIT for this cycle:
   * [0] internal, [1] translational kinetic, [2] rotational kinetic, [3] WFEXT */
  double *ecrit;
  double wfext;              /* external work of the imposed velocities (FIXVEL's WFEXT) */
  double *wpend;             /* [3*nn] each imposed DOF's DT2*DW, added at the next cycle (VEL(4,N)) */
  double *dtel_h, *dtel_q;   /* [ne] per-element time-step candidates of output cycles, or NULL */
  int32_t itypts;            /* controlling element of the last output cycle: 1 solid, 3 shell, 0 none */
  int64_t nelts;             /* its 0-based index in its family, -1 none */
  /* time history (HIST2 at the same output point as ECRIT): with ecrit set, ecrit[4..6] receive
   * ECRIT's XMOMT, YMOMT, ZMOMT (ecrit.F:221-228) and, if th is non-NULL, th[18*n ..] node n's
   * D, V, A, VR, AR, X as THNOD reads them (thnod.F:59-97) */
  double *th;
  /* EHOUR (shvis3.F:425-464, chvis3.F:299-351): if ehou_h / ehou_q are non-NULL (scratch of
   * ne_hex / ne_q4 doubles), every cycle adds to ehour the sum of each element group's EHOU, the
   * groups in order, solids then shells */
  double *ehou_h, *ehou_q;
  double ehour;
} or_model_c;

void or_cycle_model(or_model_c *m, or_time_state *ts, int nthreads);
void or_nodes_model(or_model_c *m, or_time_state *ts, double dt2t, int mstop);
void or_energy_model(const or_model_c *m, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
