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

#ifdef __cplusplus
}
#endif
#endif
