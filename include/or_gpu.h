/*
 * or_gpu.h -- C ABI of the MI355X explicit element-force / time-integration path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (OR_OK == 0); or_last_error() returns the message of the last failure of
 * the calling thread.  Host arrays use the reference's Fortran layouts:
 *   X(3,NUMNOD), V(3,NUMNOD), VR(3,NUMNOD), D(3,NUMNOD)   node-major triples
 *                                       (common_source/modules/nodal_arrays.F90:109-177)
 *   SIG(NEL,6), FOR(NEL,5), MOM(NEL,3), STRA(NEL,8), HOURG(NEL,5), EINT(NEL,2)
 *                                       component-major (elbufdef_mod.F90:739-1016)
 *   LBUF SIG(NEL,5) per through-thickness point, points outermost
 *   IXS(11,NUMELS) rows 2..9, IXC(7,NUMELC) rows 2..5   1-based node ids (element_mod.F90:1174,1178)
 *   FSKY(8,LSKY), IADS(8,*), IADC(4,*), ADSKY(NUMNOD+1)  Parith/ON skyline
 *                                       (scumu3p.F, cupdt3.F:967, asspar4.F:164-180)
 * The resident-model API takes 0-based indices (documented per call).
 *
 * Reference interfaces replaced (file:line in the OpenRadioss tree):
 *   or_model_*            the resident GPU handle of the shell offload module
 *                         (engine/source/elements/shell/coque/shell_gpu_mod.F90:279-715,
 *                         shell_gpu_driver.h:35-210: create / allocate / upload / step /
 *                         download / destroy), extended to hexes and to the nodal update.
 *   or_model_run          one RESOL cycle of the path per iteration: FORINTC -> CFORC3
 *                         (resol.F:4138, forintc.F:452), FORINT -> SFORC3 (resol.F:4225,
 *                         forint.F:1159), ASSPAR4 (resol.F:4858), the DT2 pipeline
 *                         (resol.F:6125-6129, 6352, 6494-6497), ACCELE (resol.F:6921),
 *                         BCS10 (resol.F:7322), FIXVEL (resol.F:7610), VELOCITY (resol.F:8947),
 *                         DEPLA (resol.F:9042), time advance (resol.F:8599-8613).
 *   or_sforc3_group       SFORC3 for one element group (sforc3.F:131-166) on the Isolid=1,
 *                         law-2, Lagrangian branch; writes FSKY rows like SCUMU3P (scumu3p.F:37).
 *   or_cforc3_group       CFORC3 for one element group (cforc3.F:85-113) on the Ishell=1,
 *                         law-2 branch; writes FSKY rows like CUPDT3P (cupdt3.F:967).
 *   or_model_energy       internal / kinetic energy sums of ECRIT (ecrit.F:222-275, 344).
 */
#ifndef OR_GPU_H
#define OR_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_OK 0
#define OR_ERR_ARG 1      /* invalid argument (sizes, null pointers, unsupported option) */
#define OR_ERR_HIP 2      /* HIP runtime failure */
#define OR_ERR_NODEV 3    /* no usable GPU */
#define OR_ERR_STATE 4    /* call out of order (e.g. run before upload) */
#define OR_ERR_MSTOP 5    /* an element reported a non-positive Jacobian (MSTOP=1) */

#define OR_PREC_FP64 0    /* plain method: everything in FP64 */
#define OR_PREC_SPLIT 1   /* FP32 element arithmetic/state, FP64 nodes and reductions */

/* Node code bits (bcs1.F LCOD convention per direction):
 * BCS translations 4=x 2=y 1=z, BCS rotations 32=x 16=y 8=z, imposed velocity 256=x 128=y 64=z. */
#define OR_BC_TX 4
#define OR_BC_TY 2
#define OR_BC_TZ 1
#define OR_BC_RX 32
#define OR_BC_RY 16
#define OR_BC_RZ 8
#define OR_IMP_X 256
#define OR_IMP_Y 128
#define OR_IMP_Z 64

/* Law 2 (Johnson-Cook) + /PROP/SOLID Isolid=1 + run control, resolved as the
 * Starter stores them (hm_read_mat02_jc.F90, hm_read_prop14.F:355-369). */
typedef struct {
  double rho0, bulk, shear;
  double ca, cb, cn, epmx, sigmx, cc, epdr, fisokin, asrate, m_exp;
  int32_t icc, vp, israte, iform;
  double qa, qb, qh, vns1, vns2;
  double dtfac; /* DTFAC1(1) */
  double dtmx;  /* DTMX */
} or_hex_params;

/* Law 2 for shells as SIGEPS02C reads it, the PM words CCOEF3 reads, and
 * /PROP/SHELL Ishell=1 IGTYP=1 (hm_read_mat02_jc.F90:236-290, hm_read_prop01.F:156-245). */
typedef struct {
  double rho, young, nu, shear, ssp;            /* PM(1), PM(20), PM(21), PM(22), PM(27) */
  double ca, cb, cn, epmx, ymax, cc, epdr, m_exp; /* uparam(1..7,10) */
  double asrate;                                /* strain-rate filter 2*pi*fcut (israte>0) */
  int32_t icc, vp, israte, iform;
  int32_t npt;                                  /* through-thickness points, 2..10 */
  int32_t ipla;                                 /* Iplas 0, 1 (Newton), 2 */
  int32_t ithk;                                 /* Ithick */
  int32_t ishear;                               /* GEO(37) */
  double hm, hf, hr;                            /* GEO(13..15) hourglass coefficients */
  double fsh;                                   /* GEO(38) shear factor */
  double dtfac;                                 /* DTFAC1(3) */
  double dtmin;                                 /* DTMIN1(3) */
} or_q4_params;

/* Cycle scalars (resol.F: TT, DT1, DT2, DT12, DT2OLD, NCYCLE, MSTOP).  MSTOP bits:
 * OR_MSTOP_JACOBIAN, an element reported a non-positive Jacobian (schkjabt3.F);
 * OR_MSTOP_TIME_STEP, the cycle's time step came out <= 0 (resol.F:8575, 9599).  A model
 * with MSTOP set stops cycling. */
#define OR_MSTOP_JACOBIAN 1
#define OR_MSTOP_TIME_STEP 2
typedef struct {
  double tt, dt1, dt2, dt12, dt2old;
  int64_t ncycle;
  int32_t mstop;
} or_time_state;

typedef struct or_model or_model;

/* ---- library ---- */
int or_version(void);
const char *or_last_error(void);
int or_device_count(int *count);
/* Free and total device memory in bytes (no counterpart in the reference: the serving
 * loop sizes its HBM budget with it). */
int or_device_memory(int device, int64_t *free_bytes, int64_t *total_bytes);

/* ---- resident model (upload once, step N cycles, state stays in HBM) ---- */
/* nn nodes, ne_hex 8-node hexes, ne_q4 Q4 shells with npt through-thickness points,
 * precision OR_PREC_*. */
int or_model_create(or_model **out, int device, int64_t nn, int64_t ne_hex, int64_t ne_q4, int32_t npt,
                    int precision);
int or_model_destroy(or_model *m);
int or_model_set_hex_params(or_model *m, const or_hex_params *p);
int or_model_set_q4_params(or_model *m, const or_q4_params *p);
/* x,v: X(3,nn), V(3,nn); vr: VR(3,nn) or NULL (zero); ms: (nn); in: (nn) nodal rotational
 * inertia, required when the model has shells; code: (nn) node code bits or NULL;
 * vimp: V(3,nn) imposed velocities, required when any imposed-velocity bit is set. */
int or_model_upload_nodes(or_model *m, const double *x, const double *v, const double *vr, const double *ms,
                          const double *in, const int32_t *code, const double *vimp);
/* d: D(3,nn), the displacements of a restart (or_model_upload_nodes sets D to zero).  With
 * the nodes, the element state and or_model_set_time, it restores a model from the arrays
 * the or_model_download_* calls return, as the Engine's restart file carries X, D, V and VR
 * (wrrestp.F:1188-1194) with the element buffers. */
int or_model_upload_displacement(or_model *m, const double *d);
/* conn: (ne,8) 0-based node ids; state arrays: sig SIG(NEL,6) component-major over all hexes,
 * eint, rho, vol0, pla, epsd, offg, qvis (ne each); NULL state arrays start at zero (offg at one).
 * Element arrays are always given and returned in the caller's element order.  The first
 * element upload of a family fixes its storage order on the device: tiles of 256 elements
 * ordered along a space-filling curve of the element centroids when the nodes are already
 * uploaded (the order the LDS assembly profits from), the caller's order otherwise.  Later
 * uploads may change the state, not the connectivity. */
int or_model_upload_hex(or_model *m, const int32_t *conn, const double *sig, const double *eint,
                        const double *rho, const double *vol0, const double *pla, const double *epsd,
                        const double *offg, const double *qvis);
/* conn: (ne,4) 0-based node ids (node 4 == node 3 for a triangle); thke: initial thickness;
 * GBUF: offg, thk, forc FOR(NEL,5), mom MOM(NEL,3), gstr STRA(NEL,8), hourg HOURG(NEL,5),
 * eint EINT(NEL,2), epsd; LBUF per point: sig (npt,5,ne), pla (npt,ne), epsd_ip (npt,ne).
 * NULL state arrays start at the Starter's values (OFF=1, THK=THKE, everything else zero). */
int or_model_upload_q4(or_model *m, const int32_t *conn, const double *thke, const double *offg, const double *thk,
                       const double *forc, const double *mom, const double *gstr, const double *hourg,
                       const double *eint, const double *epsd, const double *sig, const double *pla,
                       const double *epsd_ip);
/* adsky: (nn+1) 0-based CSR pointers; iads: (ne_hex,8) and iadc: (ne_q4,4) 0-based global rows.
 * The rows of a node must list its solid rows before its shell rows (FILLCNE order). */
int or_model_upload_skyline(or_model *m, const int64_t *adsky, const int64_t *iads, const int64_t *iadc);
int or_model_set_time(or_model *m, const or_time_state *ts);
int or_model_get_time(or_model *m, or_time_state *ts);
/* Run ncycles cycles asynchronously on the model's stream. */
int or_model_run(or_model *m, int64_t ncycles);
int or_model_sync(or_model *m);
/* One cycle with ECRIT's energy output (ecrit.F:193-327 at its call point resol.F:8523,
 * between ACCELE/BCS10/FIXVEL and VELOCITY): out[0] internal energy (SBILAN EINT*VOL,
 * CBILAN EINT(1)+EINT(2)), out[1] translational kinetic sum (V+DT05*A)^2*HALF*MS,
 * out[2] rotational kinetic sum (VR+DT05*AR)^2*HALF*IN, DT05 = HALF*DT1.  Synchronous. */
int or_model_run_output(or_model *m, double out[3]);
/* What ECRIT prints for one output cycle (ecrit.F:651, FORMAT 1100 at ecrit.F:704).
 * wfext: OUTPUT%TH%WFEXT, the external work of the imposed velocities since the last node
 * upload, FIXVEL's DT1*DW per imposed DOF (fixvel.F:376-378, 834), accumulated per node and
 * summed over the nodes.  itypts / nelts: the element that controls the cycle's time step
 * (ITYPTS 1 solid, 3 shell, 0 none; NELTS as a 0-based index into that family's element
 * arrays), chosen as MQVISCB and CDT3 choose it (mqviscb.F:622-630, cdt3.F:205-215): the last
 * solid with the smallest time step, unless a shell's is smaller, then the first such shell. */
typedef struct {
  double enint, encin, enrot, wfext;
  int32_t itypts, pad;
  int64_t nelts;
} or_ecrit;
/* or_model_run_output with the listing's external work and controlling element.  Synchronous. */
int or_model_run_ecrit(or_model *m, or_ecrit *out);
/* The external work of the imposed velocities so far (as in or_ecrit.wfext).  Synchronous. */
int or_model_wfext(or_model *m, double *wfext);
/* Time history (HIST2, hist2.F:299-330 and 959-999, called from sortie_main.F:1049 at ECRIT's
 * point of the cycle).  or_model_th_setup lists the entities of the /TH/NODE, /TH/BRIC and
 * /TH/SHEL groups (0-based node, hex and shell indices, each node once) and starts the EHOUR
 * accumulation (shvis3.F:464, chvis3.F:397: every cycle adds its elements' hourglass energy);
 * a later call replaces the lists and keeps EHOUR. */
int or_model_th_setup(or_model *m, int64_t n_node, const int32_t *node, int64_t n_hex, const int32_t *hex,
                      int64_t n_q4, const int32_t *q4);
/* One output cycle (as or_model_run_ecrit, whose record it fills) with the time-history values:
 * glob = XMOMT, YMOMT, ZMOMT (ecrit.F:226-228, (V+DT05*A)*MS summed over the nodes) and EHOUR;
 * node_out[n_node][18] = D, V, A, VR, AR, X as THNOD reads them (thnod.F:59-97: D and X before
 * DEPLA, V after BCS10, A after FIXVEL; VR, AR zero without rotational DOFs);
 * hex_out[n_hex][13] = OFF, SIG(6), EINT, RHO, QVIS, VOL0, PLA, EPSD of THSOL (thsol.F:189-200,
 * 511-548); q4_out[n_q4][12+NPT] = FOR(5), MOM(3), EINT(2), OFF, THK and each point's PLA of
 * THCOQ (thcoq.F:296-340, 452-520).  Synchronous. */
int or_model_run_th(or_model *m, or_ecrit *ec, double glob[4], double *node_out, double *hex_out, double *q4_out);
/* EHOUR so far; and its value set (a restart), after or_model_th_setup.  Synchronous. */
int or_model_ehour(or_model *m, double *ehour);
int or_model_set_ehour(or_model *m, double ehour);
/* Any output pointer may be NULL. */
int or_model_download_nodes(or_model *m, double *x, double *v, double *vr, double *d);
int or_model_download_hex(or_model *m, double *sig, double *eint, double *rho, double *pla,
                          double *epsd, double *offg, double *qvis);
int or_model_download_q4(or_model *m, double *offg, double *thk, double *forc, double *mom, double *gstr,
                         double *hourg, double *eint, double *epsd, double *sig, double *pla, double *epsd_ip);
/* out[0] = internal energy (hex EINT*VOL0 + shell EINT(1)+EINT(2)), out[1] = 0.5*sum MS*|V|^2,
 * out[2] = 0.5*sum IN*|VR|^2; fixed-order deterministic reductions in FP64. */
int or_model_energy(or_model *m, double out[3]);
/* Average device time per launch (ms) of the hex, Q4 and node kernels over the cycles run
 * since the last reset, measured with HIP events on the model stream. */
int or_model_kernel_times(or_model *m, int reset, double *hex_ms, double *q4_ms, double *node_ms, int64_t *cycles);
int or_model_set_timing(or_model *m, int enable);
/* Bytes of device memory held by the model. */
int or_model_footprint(const or_model *m, int64_t *bytes);
/* Diagnostic, no counterpart in the reference: the assembly plan of the uploaded skyline,
 * out = {hex prefixes, hex rest rows, shell prefixes, shell rest rows}.  A node's leading
 * rows that share the tile of its first element are summed in that tile's LDS into one
 * prefix; the other rows ("rest") are stored and summed by the node kernel. */
int or_model_assembly_stats(const or_model *m, int64_t out[4]);
/* An upper bound of the bytes a model will hold once created and uploaded, computed
 * without allocating anything (admission by HBM footprint); it counts every skyline
 * row as a stored row, where the element tiles sum most of them in LDS.  strain_rate_filter: the shells' law has
 * the strain-rate filter on (ISRATE > 0); rows_hex / rows_q4: (nn) number of hex /
 * shell skyline rows of each node (required when the family is present). */
int or_model_footprint_estimate(int64_t nn, int64_t ne_hex, int64_t ne_q4, int32_t npt, int precision,
                                int strain_rate_filter, const int32_t *rows_hex, const int32_t *rows_q4,
                                int64_t *bytes);

/* ---- group-level drop-in for SFORC3 (one group of nel <= 128 hexes) ----
 * Host arrays in the reference layouts: x X(3,numnod), v V(3,numnod), ixs IXS(11,*)
 * with the group starting at column nft+1, GBUF arrays of the group (sig SIG(nel,6)),
 * fsky FSKY(8,lsky), iads IADS(8,*) with 1-based rows (column nft+i for element i).
 * *dt2t is min-folded with the group's element time steps like MQVISCB.  dt1 = DT1. */
int or_sforc3_group(int device, const or_hex_params *p, int64_t numnod, int32_t nel, int64_t nft,
                    const double *x, const double *v, const int32_t *ixs, double *sig, double *eint,
                    double *rho, const double *vol, double *pla, double *epsd, double *off,
                    double *qvis, int64_t lsky, double *fsky, const int32_t *iads, double dt1,
                    double *dt2t, int32_t *mstop);

/* Allocation-free variant of or_sforc3_group: no counterpart in the reference.  It is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation): it calls no hipMalloc / hipFree and stages the
 * group through one file-scope device array of 139,264 bytes per device, held until the
 * process ends (the group's own nodes, at most 8 per element, its GBUF arrays and its FSKY
 * rows).  Same arguments, results and errors as or_sforc3_group; nel is at most 128.  Every
 * call on a device uses that device's one array: calls that overlap on one device race. */
int or_sforc3_group_static(int device, const or_hex_params *p, int64_t numnod, int32_t nel, int64_t nft,
                           const double *x, const double *v, const int32_t *ixs, double *sig, double *eint,
                           double *rho, const double *vol, double *pla, double *epsd, double *off,
                           double *qvis, int64_t lsky, double *fsky, const int32_t *iads, double dt1,
                           double *dt2t, int32_t *mstop);

/* ---- group-level drop-in for CFORC3 (one group of nel <= 128 Q4 shells) ----
 * x, v, vr X/V/VR(3,numnod); ixc IXC(7,*) from column nft+1; GBUF arrays of the group;
 * LBUF sig (npt,5,nel), pla (npt,nel), epsd_ip (npt,nel); fsky FSKY(8,lsky);
 * iadc IADC(4,*) 1-based.  *dt2t is min-folded with the group's time steps like CDT3. */
int or_cforc3_group(int device, const or_q4_params *p, int64_t numnod, int32_t nel, int64_t nft,
                    const double *x, const double *v, const double *vr, const int32_t *ixc, double *offg,
                    double *thk, const double *thke, double *forc, double *mom, double *gstr, double *hourg,
                    double *eint, double *epsd, double *sig, double *pla, double *epsd_ip, int64_t lsky,
                    double *fsky, const int32_t *iadc, double dt1, double *dt2t);

/* ---- nodal half of a cycle, drop-in with the reference host arrays ----
 * ASSPAR4 (asspar4.F:164-180) over FSKY(8,lsky) rows listed by ADSKY(numnod+1) (1-based
 * CSR, FILLCNE order), the DT2 pipeline from the element minimum dt2t (resol.F:2721-2722,
 * 6125-6129, 6352, 6494-6497), ACCELE (accele.F), BCS10 / FIXVEL (bcs10.F, fixvel.F:366-379)
 * with the node code bits and vimp(3,numnod), VELOCITY (velocity.F:58-86), DEPLA
 * (displacement.F:108-118) and the time advance (resol.F:8599-8613) of *ts.
 * x, v, d, vr: X/V/D/VR(3,numnod) updated in place (vr and in NULL without rotations);
 * stifn, stifr: STIFN/STIFR(numnod) from FSKY words 7, 8, or NULL. */
int or_nodal_update(int device, int64_t numnod, int64_t lsky, const int64_t *adsky, const double *fsky,
                    const double *ms, const double *in, const int32_t *code, const double *vimp, double *x,
                    double *v, double *vr, double *d, double *stifn, double *stifr, double dt2t, double dtmx,
                    or_time_state *ts);

/* ---- batches: many resident models advanced together ----
 * No counterpart in the reference, which runs one deck per process: a batch advances
 * every member model by its own explicit cycles (its own DT2 pipeline, time step and
 * state) in one launch per kernel and precision class per cycle, so that many small
 * models fill one GPU.  A member stops after the cycle whose time reaches its t_end
 * (or when an element raises MSTOP) and then keeps its state.  Members stay owned by
 * the caller; their or_model_download_* / get_time / energy calls work between
 * batch calls.  or_model_run is refused while a model belongs to a batch. */
typedef struct or_batch or_batch;
int or_batch_create(or_batch **out, int device);
int or_batch_destroy(or_batch *b);   /* removes the members, does not destroy them */
int or_batch_add(or_batch *b, or_model *m, double t_end);
int or_batch_remove(or_batch *b, or_model *m);
/* Launch ncycles cycles of every member still running (asynchronous). */
int or_batch_run(or_batch *b, int64_t ncycles);
int or_batch_sync(or_batch *b);
/* Number of members still running; synchronises, and drops stopped members from the
 * launches of later cycles. */
int or_batch_active(or_batch *b, int32_t *nactive);

/* ---- launch variants of the element kernels and their autotuning ----
 * No counterpart in the reference.  Variant v (0..3) of the hex (family 0) or shell
 * (family 1) kernel sets the register budget it is compiled for (two waves per SIMD for
 * even v, three for odd v) and the compiler's instruction scheduler (default for v < 2,
 * ILP-first for v >= 2); every workgroup is one tile of 256 elements.  All variants
 * compute bit-identical results.  The autotuners time cycles_per_variant real
 * cycles per candidate (the models advance meanwhile) and keep the fastest:
 * chosen[family] for a model, chosen[2*precision + family] for a batch. */
int or_model_set_variant(or_model *m, int family, int variant);
int or_model_autotune(or_model *m, int32_t cycles_per_variant, int32_t chosen[2]);
int or_batch_autotune(or_batch *b, int32_t cycles_per_variant, int32_t chosen[4]);

/* Diagnostic, no counterpart in the reference: evaluates on the device the FP64
 * elementary functions the element kernels use in place of the host libm's
 * (fn 0: log(x), 1: exp(x), 2: pow(x, y)), for n arguments, 0 <= n <= 2^30.
 * They are correctly rounded except within ~2^-17 ulp of a rounding boundary. */
int or_math_eval(int device, int fn, int64_t n, const double *x, const double *y, double *out);

#ifdef __cplusplus
}
#endif
#endif
