/*
 * qws_mi355x.h -- C ABI of libqws_mi355x.so, the MI355X (gfx950) Wilson-clover
 * domain-decomposed solver.
 *
 * Two interfaces:
 *
 *  1. QWS drop-in.  The Fortran-callable symbols (trailing underscore, scalars by
 *     pointer) of QWS's DD solver path, with QWS's host data layout (VLEND = 8,
 *     VLENS = 16 AoSoA, qws.h:111-192) and QWS's exported globals.  A caller
 *     written against QWS (its test driver main.cc, or BQCD through
 *     qws_bqcd.cc) relinks against this library unchanged.  Arguments, argument
 *     meaning and error behaviour follow the reference; each declaration cites
 *     the reference interface it replaces.  Single rank: npe must be (1,1,1,1).
 *
 *  2. Device-resident interface (qwsg_*).  Fields and spinors live in HBM in the
 *     library's blocked site-major layout; nothing crosses PCIe inside a solve.
 *     Plain pointers and sizes only; every function returns 0 on success or a
 *     nonzero error code, with qwsg_last_error() describing the failure.
 */
#ifndef QWS_MI355X_H
#define QWS_MI355X_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- QWS host types
 * Same memory layout as qws.h:111-192 with VLEND = 8, VLENS = 16 and 18-real links
 * (SU3_RECONSTRUCT_D/S = 18, qws.h:60-61). */
#define QWS_VLEND 8
#define QWS_VLENS 16
typedef struct { double c[3][4][2][QWS_VLEND]; } scd_t;   /* qws.h:134-139 */
typedef struct { float c[3][4][2][QWS_VLENS]; } scs_t;    /* qws.h:176-182 */
typedef struct { double c[3][3][2][QWS_VLEND]; } glud_t;  /* qws.h:121-127 */
typedef struct { float c[3][3][2][QWS_VLENS]; } glus_t;   /* qws.h:160-168 */
typedef struct { double c[2][36][QWS_VLEND]; } clvd_t;    /* qws.h:129-132 */
typedef struct { float c[2][36][QWS_VLENS]; } clvs_t;     /* qws.h:170-174 */

/* ---------------------------------------------------------------- QWS drop-in: globals
 * Host mirrors of QWS's exported data symbols (qws.cc:74-124).  qws_init_
 * allocates the four field arrays; a caller may write them directly, as the
 * reference driver does (main.cc:439-516).  The device copy is refreshed on the
 * next solver/operator call after qws_init_, after any qws_load*_ call, or after
 * qws_mi355x_fields_modified_(). */
extern glud_t* glud;
extern glus_t* glus;
extern clvd_t* clvd;
extern clvs_t* clvs;
extern double kappa, kappa2, mkappa;
extern int nx, ny, nz, nt, nxh, nxd, nxs, vold, vols;
extern int npe[4];
extern double fbc[4][2];
extern int rank, domain_e, domain_o;
/* qws.cc:124 -- QWS's per-rank output file (opened there only for per-rank timing output) */
extern FILE* para_outputfile;

/* qws.cc:130 -- geometry, field allocation, t boundary (fbc_f[3] = -1: antiperiodic).
 * block_size is ignored, as in the reference; the SAP domains are the two x-halves
 * unless qws_mi355x_set_sap_blocks_ is called. */
void qws_init_(int* lx, int* ly, int* lz, int* lt, int* npe_f, int* fbc_f, int* pce_f, int* pco_f, int* block_size);
/* qws.cc:319 */
void qws_finalize_(void);

/* qws_bqcd.cc:125 -- BQCD-ordered links; sets kappa.  BQCD's std_xyzt2i_ and e_o_
 * (site index within its parity, 1-based, and the parity) must be linked or loaded
 * globally before any qws_load / qws_store call. */
void qws_loadg_dd_(double* u, int* volh_tot, double* in_kappa);
/* qws_bqcd.cc:284 -- BQCD-ordered clover inverse (72 reals per site) */
void qws_loadc_dd_(double* c, int* volh);
/* qws_bqcd.cc:500,536,572,608 -- BQCD even/odd spinors <-> DD layout */
void qws_loadfd_dd_(scd_t* out, double* ine, double* ino);
void qws_storefd_dd_(double* oute, double* outo, scd_t* in);
void qws_loadfs_dd_(scs_t* out, double* ine, double* ino);
void qws_storefs_dd_(double* oute, double* outo, scs_t* in);
/* qws_bqcd.cc:66,184,384,413,442,471 -- the same BQCD arrays into the even/odd reading
 * of the host arrays (x-compressed sites, parity slot = e_o_): links and clover inverse
 * of both parities; spinors of the even sites only (a field of vold / vols vectors). */
void qws_loadg_bqcd_(double* u, int* volh_tot, double* in_kappa);
void qws_loadc_bqcd_(double* c, int* volh);
void qws_loadfd_bqcd_(scd_t* out, double* in);
void qws_loadfs_bqcd_(scs_t* out, double* in);
void qws_storefd_bqcd_(double* out, scd_t* in);
void qws_storefs_bqcd_(double* out, scs_t* in);

/* bicgstab_dd_mix.cc:105 -- FP64 outer BiCGStab with FP32 SAP-preconditioned inner solve.
 * x: initial guess in, solution out.  tol < 0: run maxiter-1 outer iterations. */
void bicgstab_dd_mix_(scd_t* x, scd_t* b, double* tol, int* conviter, int* maxiter, double* tol_s, int* maxiter_s,
                      int* nsap, int* nm);
/* Not in the reference: bicgstab_dd_mix_ for callers that need to know the outcome.
 * Same arguments and results, plus true_residual = |b - D x| / |b| recomputed in FP64 on
 * the GPU (-1 if the solve did not run).  Returns 0 when the recursive residual met tol
 * (or tol < 0), 1 when maxiter ran out, 2 when the residual is not finite (breakdown),
 * -1 before qws_init_, -2 for invalid arguments, -3 for a device or runtime failure;
 * errors are reported on stderr and never end the process. */
int bicgstab_dd_mix_status_(scd_t* x, scd_t* b, double* tol, int* conviter, int* maxiter, double* tol_s,
                            int* maxiter_s, int* nsap, int* nm, double* true_residual);
/* Not in the reference: bicgstab_dd_mix2_hf_ (the FP16-storage SAP method) with the same
 * return codes and true residual as bicgstab_dd_mix_status_; *conviter as
 * bicgstab_dd_mix2_hf_ reports it (set only on convergence). */
int bicgstab_dd_mix2_hf_status_(scd_t* x, const scd_t* b, const double* tol, int* conviter, const int* maxiter,
                                const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm,
                                double* true_residual);
/* bicgstab_precdd_s.cc:85 -- FP32 right-preconditioned BiCGStab on D M_SAP, x0 = 0 */
void bicgstab_precdd_s_(scs_t* x, scs_t* b, double* tol, int* conviter, int* maxiter, int* nsap, int* nm);
/* qws.cc:776 -- out = D M_SAP in */
void prec_ddd_s_(scs_t* out, scs_t* in, int* nsap, int* nm);
/* qws.cc:929 -- out = M_SAP in */
void prec_s_(scs_t* out, scs_t* in, int* nsap, int* nm);
/* Half-precision variant (qws_h.h:60-62, bicgstab_dd_mix2_hf.cc:110, bicgstab_precdd_s_h.cc:83).
 * qws_h_init_ makes the FP16 copies of the FP32 links and clover inverse;
 * assign_mult_mSAP_s_h_ is M_SAP with FP16-stored iterates and fields (nsap - 1
 * sweeps plus the final block solves, FP32 arithmetic); assign_mult_wd_s_ is the
 * FP32 D; bicgstab_precdd_s_h_ is the inner BiCGStab whose x accumulates M_SAP p
 * directly; bicgstab_dd_mix2_hf_ is the FP64 outer solve on top of it. */
void qws_h_init_(const glus_t* u, const clvs_t* c);
/* qws_bqcd.cc:647 -- qws_h_init_(glus, clvs) */
void bqcd_qws_h_init_(void);
void assign_mult_mSAP_s_h_(scs_t* xs, const scs_t* bs, const int* nsap, const int* nm);
void assign_mult_wd_s_(scs_t* x, const scs_t* b);
void bicgstab_precdd_s_h_(scs_t* x, const scs_t* b, const double* tol, int* conviter, const int* maxiter,
                          const int* nsap, const int* nm);
void bicgstab_dd_mix2_hf_(scd_t* x, const scd_t* b, const double* tol, int* conviter, const int* maxiter,
                          const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm);
/* static_solver.cc:141 -- Jacobi block solve on one domain (x, b point at that domain) */
void jinv_ddd_in_s_(scs_t* x, scs_t* b, int* domain, int* maxiter);
/* qws.cc:625,641 -- full operator D = 1 - kappa T^-1 H */
void ddd_d_(scd_t* out, scd_t* in);
void ddd_s_(scs_t* out, scs_t* in);
/* ddd_in_d.cc:66, ddd_in_s.cc:376 -- domain-local operator (out, in point at that domain) */
void ddd_in_d_(scd_t* out, scd_t* in, int* domain);
void ddd_in_s_(scs_t* out, scs_t* in, int* domain);
/* ddd_out_d.cc:84,164, ddd_out_s.cc:639,661 -- face terms of a domain.  pre records
 * the faces of `in`; pos adds factor * T^-1 H_face (FP64: factor -kappa) to out,
 * which points at that domain. */
void ddd_out_pre_d_(scd_t* in, int* domain);
void ddd_out_pos_d_(scd_t* out, scd_t* in, int* domain);
void ddd_out_pre_s_(scs_t* in, int* domain);
void ddd_out_pos_s_(scs_t* out, scs_t* in, int* domain, float factor);
/* qws.cc:390,429 -- clover inverse multiply on one parity/domain block of vold sites */
void clvd_vm_(int* pe, scd_t* inout);
void clvd_vm2_(int* pe, scd_t* out, scd_t* in);
/* Even/odd family (main.cc:554-721 lists them).  The host arrays glud/glus, clvd/clvs
 * are read as two parity halves of x-compressed sites (qws_bqcd.cc:82-85): a field
 * of parity p holds vold (FP64) or vols (FP32) vectors, vector i = xh/VLEN +
 * nxd*(y + ny*(z + nz*t)), lane xh % VLEN.  out has parity *pe, in parity *po.
 * "_in_": the hops that stay inside the lattice; "_out_pre_/_pos_": the hops that
 * wrap round it (pre records in, pos adds factor times them to out); "dee_": T^-1
 * applied; "dag": H^+; "one_minus_": out = in0 + factor (...).
 * deo_in_d.cc:65,191,319; deo_out_d.cc:88,156,220; deo_dag_in.cc:64,190,317;
 * deo_dag_out.cc:88,156,220; qws.cc:518-576 (composites, mtilde with pce/pco). */
void deo_in_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void dee_deo_in_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void deo_dag_in_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void dee_deo_dag_in_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void one_minus_dee_deo_in_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0, double factor);
void one_minus_deo_dag_in_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0, double factor);
void deo_out_pre_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void deo_dag_out_pre_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void deo_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor);
void deo_dag_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor);
void dee_deo_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor);
void dee_deo_dag_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor);
void deo_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void dee_deo_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void one_minus_dee_deo_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0);
void deo_dag_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void dee_deo_dag_vm_(int* pe, int* po, scd_t* out, scd_t* in);
void one_minus_deo_dag_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0);
void mtilde_vm_(scd_t* out, scd_t* in);
void mtilde_dag_vm_(scd_t* out, scd_t* in);
void mtdagmt_vm_(scd_t* out, scd_t* in);
/* bicgstab.cc:65 -- FP64 BiCGStab on mtilde_vm_ (x: initial guess in, solution out;
 * stops at sqrt(|r|^2/|b|^2) < 1e-14 or after maxiter-1 iterations) */
void bicgstab_vm_(scd_t* x, scd_t* b, int* conviter, int* maxiter);
/* cg.cc:65 -- CG on mtdagmt_vm_ (x: initial guess in, solution out; stops at
 * sqrt(|r|^2/|b|^2) < 1e-14 or after maxiter-1 iterations).  *conviter receives the
 * iterations run (the reference does not set it). */
void cg_vm_(scd_t* x, scd_t* b, int* conviter, int* maxiter);
/* mcg.cc:65 -- multi-shift CG: for k < *n, xx[k*vold ...] (vold sites each) solves
 * (mtdagmt_vm_ + sigma[k]) x = b, from zero, with the stopping rule of cg_vm_ on the
 * unshifted residual.  1 <= *n <= 64. */
void mcg_vm_(scd_t* xx, scd_t* b, int* n, double* sigma, int* maxiter);
/* FP32: deo_in_s.cc:65,191,317; deo_out_s.cc:93,155,229; qws.cc:576-605; clover_s.cc (clv_s_) */
void deo_in_s_(int* pe, int* po, scs_t* out, scs_t* in);
void dee_deo_in_s_(int* pe, int* po, scs_t* out, scs_t* in);
void one_minus_dee_deo_in_s_(int* pe, int* po, scs_t* out, scs_t* in, scs_t* in0, float factor);
void deo_out_pre_s_(const int* pe, const int* po, const scs_t* out, const scs_t* in);
void deo_out_pos_s_(const int* pe, const int* po, const scs_t* out, const scs_t* in, float factor);
void dee_deo_out_pos_s_(int* pe, int* po, scs_t* out, scs_t* in, float factor);
void clv_s_(int* pe, scs_t* inout);
void deo_s_(int* pe, int* po, scs_t* out, scs_t* in);
void dee_deo_s_(int* pe, int* po, scs_t* out, scs_t* in);
void one_minus_dee_deo_s_(int* pe, int* po, scs_t* out, scs_t* in, scs_t* in0);
void mtilde_s_(scs_t* out, scs_t* in);
/* reordering.cc:141,108 -- FP64 <-> FP32 with normalisation */
void fermi_reorder_d2s_dd_renorm_(scs_t* out, const scd_t* in, double* dnorm);
void fermi_reorder_s2d_dd_renorm_(scd_t* out, const scs_t* in, const double* dnorm);
/* reordering.cc:83,59 -- FP64 <-> FP32 without normalisation */
void fermi_reorder_d2s_dd_(scs_t* out, const scd_t* in);
void fermi_reorder_s2d_dd_(scd_t* out, const scs_t* in);
/* ddd_out_s.cc:96,100 -- the face terms without per-call timers (the same functions here) */
void ddd_out_pre_s_no_timer_(scs_t* in, int* domain);
void ddd_out_pos_s_no_timer_(scs_t* out, scs_t* in, int* domain, float factor);
/* qws.cc:445-500 -- driver helpers: clear size vectors; print "%30s : %24.14e" of |field|^2 */
void zero_scd_field(scd_t* inout, int size);
void zero_scs_field(scs_t* inout, int size);
void print_scdnorm2(const char* a, scd_t* in, int size);
void print_scsnorm2(const char* a, scs_t* in, int size);
/* timing.c:171,185 -- wall clock in seconds; the timing report (this library keeps no
 * per-function timers: check_timing_(NULL) and print_timing_ say so) */
double get_clock(void);
void check_timing_(const char* name);
void print_timing_(void);
/* qws.cc:634 -- D on one domain (face terms with factor mkappa) into out + vols*domain */
void ddd_eo_s_(scs_t* out, scs_t* in, int* domain);
/* bicgstab_dd_s.cc:67, bicgstab_dd_d.cc:64 -- plain BiCGStab on ddd_s_ / ddd_d_ (x: initial
 * guess in, solution out; stops at sqrt(|r|^2/|b|^2) < 1e-6 (FP32) / 1e-14 (FP64) or
 * after maxiter-1 iterations; *conviter = the last iteration).  Reductions in FP64. */
void bicgstab_dd_s_(scs_t* x, scs_t* b, int* conviter, int* maxiter);
void bicgstab_dd_d_(scd_t* x, scd_t* b, int* conviter, int* maxiter);
/* bicgstab_precdd_s.cc:561 -- prints "@C iter err", err = |b - prec_ddd_s_(x)| / sqrt(bnorm) */
void check_residual_s_(const scs_t* x, const scs_t* b, const int* iter, const float* bnorm, const int* nsap,
                       const int* nm);

/* No counterpart in the reference: mark the host field arrays as modified, so
 * that the next call re-uploads them; and choose SAP domains finer than the two
 * x-halves (extents must divide nx/2, ny, nz, nt).  For callers that write
 * glud/glus/clvd/clvs themselves or want a GPU-sized SAP block. */
void qws_mi355x_fields_modified_(void);
/* No counterpart in the reference: upload glud/glus/clvd/clvs to the GPU now (the
 * explicit form of qws_mi355x_fields_modified_(); exits like QWS's fatal errors when
 * qws_init_ has not been called). */
void qws_gpu_sync_fields_(void);
void qws_mi355x_set_sap_blocks_(int* block_extents);
/* No counterpart as a call in the reference (QWS chooses it at compile time: its
 * default SU3_RECONSTRUCT_S = 18 at qws.h:61, the 12-real layout at qws.h:164-166 is
 * compiled out): *on != 0 stores the FP32 links on the GPU as
 * their first two rows and rebuilds row 2 as conj(row0 x row1) -- valid only for SU(3)
 * links; the host arrays keep all 18 reals.  Halves the LDS of the fused SAP kernel (two
 * workgroups per CU).  Default off (QWS's own driver's random links are not SU(3)). */
void qws_mi355x_set_links12_(int* on);

/* ---------------------------------------------------------------- device-resident interface */
typedef struct qwsg_lattice qwsg_lattice;

enum { QWSG_F64 = 0, QWSG_F32 = 1, QWSG_F16 = 2 };
enum { QWSG_HOPS_ALL = 0, QWSG_HOPS_INSIDE = 1, QWSG_HOPS_FACE = 2 };

typedef struct {
  int nsap, nm;            /* SAP sweeps and Jacobi steps per block solve (each <= 10) */
  double tol;              /* outer: recursive residual sqrt(|r|^2/|b|^2) target (<0: fixed count) */
  int maxiter;             /* outer: iterations run are 1..maxiter-1 (bicgstab_dd_mix.cc:165) */
  double tol_s;            /* inner FP32 tolerance (<0: fixed count) */
  int maxiter_s;           /* inner iterations */
  double true_tol;         /* > 0: restart until the FP64 true residual |b-Dx|/|b| <= true_tol */
  int max_restarts;        /* restarts allowed for true_tol */
  int sap_precision;       /* QWSG_F32, or QWSG_F16: the method of bicgstab_dd_mix2_hf_ (FP16-stored
                              SAP with nsap - 1 sweeps, FP32 D, inner x accumulates M_SAP p) */
} qwsg_solve_params;

typedef struct {
  int outer_iters;         /* outer BiCGStab iterations, summed over restarts */
  int inner_iters;         /* inner BiCGStab iterations, summed */
  int restarts;
  double true_residual;    /* |b - D x| / |b| recomputed in FP64 on the GPU */
  double seconds;          /* wall time of the solve */
  int sap_escalated;       /* sap_precision = QWSG_F16 and tol_s > 0: an inner solve on the FP16-stored
                              SAP ended with a relative residual >= 1 (this happens near the critical
                              kappa), and the solve went on with the FP32 SAP from there, with the
                              same number of sweeps (nsap - 1) */
} qwsg_solve_stats;

const char* qwsg_last_error(void);
int qwsg_device_count(int* n);
/* Measured HBM bandwidth (GB/s, reads + writes) of a streaming kernel that sums
 * nin arrays of `bytes` each into one (nin = 1: a copy), over reps launches; the
 * practical bandwidth ceiling the operator kernels are compared with.  Allocates
 * (nin + 1) * bytes of device memory for the call.  Measurement only; no
 * counterpart in the reference. */
int qwsg_stream_bandwidth(long long bytes, int nin, int reps, double* gbs);
/* The same with a single input stream: each 4 KB output chunk is the sum of nin
 * consecutive 4 KB chunks of one input array of nin * bytes (a read-mostly kernel
 * whose workgroups each read one contiguous region).  Measurement only; no
 * counterpart in the reference. */
int qwsg_stream_bandwidth_one(long long bytes, int nin, int reps, double* gbs);
/* Measurement only, no counterpart in the reference: the multi-RHS clover product (the
 * two Hermitian 6x6 blocks of the clover inverse, clover_def.h:88-284, applied to 16
 * right-hand sides per site, FP16 storage, FP32 accumulation, nrep products on data in
 * registers) on V synthetic sites, once on the vector ALUs and once on the matrix cores
 * (v_mfma_f32_16x16x16f16, the sites' 16 RHS as the N dimension): ms[0] VALU and ms[1]
 * MFMA device time per launch, averaged over reps launches; *maxdiff = their largest
 * difference relative to the largest VALU value. */
int qwsg_time_clover_rhs(long long V, int nrep, int reps, double* ms, double* maxdiff);

/* L: lattice extents, B: SAP block extents (must divide L; each of the
 * y/z/t block counts even or 1, x block count even), tbc: +1 periodic / -1
 * antiperiodic in t, links32: 18 (all 9 complex entries) or 12 (two rows;
 * third row rebuilt, valid for SU(3) links only). */
int qwsg_create(const int* L, const int* B, int tbc, int links32, qwsg_lattice** out);
int qwsg_destroy(qwsg_lattice* lat);
int qwsg_set_kappa(qwsg_lattice* lat, double kappa);
/* SAP kernels: 1 (default) = one fused launch per half-sweep where the geometry
 * allows it (proper colouring, 256-site blocks), 0 = one launch per operator step */
int qwsg_set_sap_fused(qwsg_lattice* lat, int on);
/* FP32 fused SAP: 1 (default) = each face update reads a face buffer that the
 * previous half-sweep wrote (projected, for backward hops link-multiplied, face
 * spinors of the other colour); 0 = it gathers the other colour's spinors and links.
 * Same results bit for bit.  No counterpart in the reference. */
int qwsg_set_sap_face_buffer(qwsg_lattice* lat, int on);
/* FP32 inner BiCGStab with the fused SAP: 1 (default) = its direction update
 * p = r + beta (p - omega q) runs inside the first SAP sweep of the next preconditioner
 * application; 0 = a separate pass.  Same results bit for bit.  No counterpart in the
 * reference. */
int qwsg_set_sap_source_update(qwsg_lattice* lat, int on);
int qwsg_volume(qwsg_lattice* lat, long long* V);
/* host plain layout: links [4][V][3][3][2] (QWS c[i][j] of the link x -> x+mu),
 * clover inverse [V][2][36] (QWS packing), site x + Lx*(y + Ly*(z + Lz*t)) */
int qwsg_load_gauge(qwsg_lattice* lat, const double* U);
int qwsg_load_clover(qwsg_lattice* lat, const double* C);
/* FP32 fields rounded from separate FP32 host arrays (QWS's glus/clvs) */
int qwsg_load_gauge_f32(qwsg_lattice* lat, const float* U);
int qwsg_load_clover_f32(qwsg_lattice* lat, const float* C);

/* No counterpart in the reference (QWS receives its fields from the caller):
 * generated inputs.  Random numbers are Philox4x32-10 with counter {site, (mu<<8)|k,
 * purpose, 0} and key = seed, so a field is the same on any launch shape.
 *   gauge : U = exp(i eps H), H = (A + A^+)/2 - tr/3, A complex Gaussian (FP64)
 *   clover: T = 1 + i kappa csw sum_{mu<nu} sigma_{mu nu} F_{mu nu} (clover leaves),
 *           each chirality block inverted in FP64, stored as QWS's 1/2 T^-1
 *           (uses the lattice's current kappa)
 *   spinor: Gaussian (re, im ~ N(0,1)) or a point source at site (x,y,z,t) */
int qwsg_generate_gauge(qwsg_lattice* lat, double eps, unsigned long long seed);
int qwsg_compute_clover(qwsg_lattice* lat, double csw);
int qwsg_gauge_download(qwsg_lattice* lat, double* U);
int qwsg_clover_download(qwsg_lattice* lat, double* C);
int qwsg_spinor_gaussian(qwsg_lattice* lat, int prec, void* dev, unsigned long long seed);
int qwsg_spinor_point(qwsg_lattice* lat, int prec, void* dev, const int* site, int spin, int colour);

/* spinors: device buffers of V sites in the given precision */
int qwsg_spinor_alloc(qwsg_lattice* lat, int prec, void** dev);
int qwsg_spinor_free(void* dev);
int qwsg_spinor_upload(qwsg_lattice* lat, int prec, void* dev, const double* host_plain);
int qwsg_spinor_download(qwsg_lattice* lat, int prec, const void* dev, double* host_plain);
int qwsg_spinor_zero(qwsg_lattice* lat, int prec, void* dev);

/* out = in - kappa T^-1 H_hops in on the sites of blocks of `colour` (-1: all) */
int qwsg_dslash(qwsg_lattice* lat, int prec, void* out, const void* in, int hops, int colour);
/* out += factor T^-1 H_face in on the sites of blocks of `colour` (-1: all) */
int qwsg_face(qwsg_lattice* lat, int prec, void* out, const void* in, double factor, int colour);
int qwsg_clover(qwsg_lattice* lat, int prec, void* out, const void* in);
int qwsg_jinv(qwsg_lattice* lat, int prec, void* x, const void* b, int colour, int nm);
int qwsg_prec_ddd(qwsg_lattice* lat, int prec, void* out, const void* in, int nsap, int nm);
int qwsg_prec_s(qwsg_lattice* lat, int prec, void* out, const void* in, int nsap, int nm);
/* No counterpart in the reference: qwsg_prec_ddd in FP32 (out = D M_SAP in) for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops that
 * must not call hipMalloc).  It allocates nothing: its work buffers (the SAP iterates
 * and the face buffer, 288 B per site with 4^4 blocks) live in a fixed device array of
 * QWSG_NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the process; a lattice
 * that needs more is rejected (return 1).  Results are bit-identical to qwsg_prec_ddd's.
 * Calls that overlap on one GPU share that array and race. */
#define QWSG_NOALLOC_SCRATCH_BYTES (192u << 20)
int qwsg_prec_ddd_noalloc(qwsg_lattice* lat, void* out, const void* in, int nsap, int nm);
int qwsg_bicgstab_precdd(qwsg_lattice* lat, int prec, void* x, const void* b, double tol, int maxiter, int nsap,
                         int nm, int* iters);
/* prec_ddd_s_h.h:63 -- out = M_SAP in with the FP16-stored SAP (FP32 vectors in and out) */
int qwsg_msap_h(qwsg_lattice* lat, void* out, const void* in, int nsap, int nm);
/* bicgstab_precdd_s_h.cc:83 -- inner solve D x = b with the FP16-stored SAP, x0 = 0 (FP32 vectors) */
int qwsg_bicgstab_precdd_h(qwsg_lattice* lat, void* x, const void* b, double tol, int maxiter, int nsap, int nm,
                           int* iters);
/* bicgstab_dd_d.cc:64 / bicgstab_dd_s.cc:67 -- plain (unpreconditioned) BiCGStab on D
 * with FP64 or FP32 vectors (FP64 reductions), x holds the initial guess; stops when
 * sqrt(|r|^2/|b|^2) < tol or after maxiter-1 iterations and returns the last one in
 * *iters.  With FP64 vectors this is the all-FP64 solve that mixed-precision results
 * are compared with. */
int qwsg_bicgstab_dd(qwsg_lattice* lat, int prec, void* x, const void* b, int maxiter, double tol, int* iters);
/* FP64 outer solve; x (FP64) holds the initial guess */
int qwsg_solve(qwsg_lattice* lat, void* x, const void* b, const qwsg_solve_params* p, qwsg_solve_stats* st);
/* No counterpart in the reference (it solves one source per call): n independent
 * right-hand sides (1..64, FP64 device spinors; x[k] holds the initial guess)
 * solved together.  Every kernel runs over all unconverged right-hand sides, so
 * the links and clover are read once per batch; each right-hand side keeps its own
 * recurrence and its own deterministic reductions, so its result is bit-identical
 * to solving it alone.  st receives n entries. */
int qwsg_solve_batch(qwsg_lattice* lat, int n, void* const* x, const void* const* b, const qwsg_solve_params* p,
                     qwsg_solve_stats* st);
/* No counterpart in the reference: solves on nlat lattices of any volumes, fields and
 * kappas at once (the SAP runs as the reference's bicgstab_dd_mix_ on each,
 * bicgstab_dd_mix.cc:105-388).  Lattice i solves nrhs[i] right-hand sides (1..64);
 * x, b and st hold them lattice after lattice (x holds the initial guesses).  The
 * lattices advance in lockstep on the first lattice's stream: every SAP half-sweep
 * and prec_ddd_s_ face term is one launch over the blocks of all of them (where
 * their block volume and link storage agree), the vector updates and reductions run
 * per lattice, so each result equals the lattice's own qwsg_solve_batch bit for bit. */
int qwsg_solve_multi(int nlat, qwsg_lattice* const* lats, const int* nrhs, void* const* x, const void* const* b,
                     const qwsg_solve_params* p, qwsg_solve_stats* st);
/* Even/odd family on x-compressed parity fields of V/2 sites h = xh + (Lx/2)(y + Ly(z + Lz t))
 * (host layout: spinor [V/2][24]; links [2][4][V/2][18], link from site h of parity p;
 * clover [2][V/2][72]).  Fields come from host arrays or from the lattice's own
 * fields (all extents even; parity-p site h is x = 2 xh + (p + y + z + t) % 2).
 * qwsg_deo: out (parity pe) from in (parity 1-pe); set 0 all hops, 1 hops inside the
 * lattice, 2 hops that wrap (t edge with 2*tbc), 3 none (in[h]); dag: H^+; clv: T^-1
 * applied; mode 0 out = acc, 1 out += fac acc, 2 out = in0 + fac acc
 * (deo_*_in/_out_*, dee_*, one_minus_*).  qwsg_mtilde: which 0 mtilde =
 * 1 - kappa^2 T^-1 H T^-1 H on parity pce, 1 its adjoint, 2 mtilde^+ mtilde
 * (qws.cc:551-576). */
int qwsg_eo_load_gauge(qwsg_lattice* lat, const double* U);
int qwsg_eo_load_clover(qwsg_lattice* lat, const double* C);
int qwsg_eo_from_lattice(qwsg_lattice* lat);
int qwsg_eo_spinor_alloc(qwsg_lattice* lat, int prec, void** dev);
int qwsg_eo_spinor_upload(qwsg_lattice* lat, int prec, void* dev, const double* host);
int qwsg_eo_spinor_download(qwsg_lattice* lat, int prec, const void* dev, double* host);
int qwsg_deo(qwsg_lattice* lat, int prec, int pe, int set, int dag, int clv, int mode, double fac, void* out,
             const void* in, const void* in0);
int qwsg_mtilde(qwsg_lattice* lat, int prec, int pce, int which, void* out, const void* in);
/* bicgstab.cc:65 -- FP64 BiCGStab on mtilde (parity pce) from the initial guess in x */
int qwsg_eo_bicgstab(qwsg_lattice* lat, void* x, const void* b, int maxiter, double tol, int pce, int* iters);
/* cg.cc:65 -- FP64 CG on mtilde^+ mtilde (parity pce) from the initial guess in x */
int qwsg_eo_cg(qwsg_lattice* lat, void* x, const void* b, int maxiter, double tol, int pce, int* iters);
/* mcg.cc:65 -- FP64 multi-shift CG: the parity field xs[k] (k < nshift, device
 * pointers) solves (mtilde^+ mtilde + sigma[k]) x = b from zero */
int qwsg_eo_mcg(qwsg_lattice* lat, void* const* xs, const void* b, int nshift, const double* sigma, int maxiter,
                double tol, int pce, int* iters);
int qwsg_norm2(qwsg_lattice* lat, int prec, const void* a, double* out);
/* No counterpart in the reference: an order-independent 64-bit digest of a device
 * spinor's bytes, the digest a served request reports for its solutions (XOR over the
 * solutions of their digests); equal for bit-identical spinors. */
int qwsg_spinor_digest(qwsg_lattice* lat, int prec, const void* a, unsigned long long* out);

/* No counterpart in the reference (it solves one source per call, serially): a solve
 * server on one GPU with continuous batching.  The gauge fields stay resident; each
 * request names a field, a kappa and 1..64 Gaussian sources (source k uses seed*1000 +
 * k).  qwsg_server_run replays a trace in real time: a request is released at its
 * arrival time and admitted in arrival order when fewer than `slots` (1..16) requests
 * run and a pooled slot for it (a lattice sharing the field's links, the clover inverse
 * for its kappa, its solver vectors) is free or fits `mem_budget` bytes (0: 90% of the
 * free HBM).  All admitted requests advance together, one outer iteration at a time,
 * every SAP half-sweep one launch over the blocks of all of them; a request that
 * arrives meanwhile joins at the next outer iteration.  Slots are kept for later
 * requests (nothing is allocated once the pool has warmed up).  Results carry arrival,
 * admission and finish times (seconds from the start of the trace), iteration counts,
 * the largest FP64 true residual and an order-independent digest of the solutions,
 * which is the same whether the request runs alone or among others. */
typedef struct qwsg_server qwsg_server;
typedef struct {
  int L[4], B[4];
  double eps;                /* links exp(i eps H) */
  unsigned long long seed;
} qwsg_field_spec;
typedef struct {
  int field;
  double kappa;
  int nrhs;
  unsigned long long seed;
  double arrival;
} qwsg_request;
typedef struct {
  double arrival, admit, finish;
  int outer_iters, inner_iters, restarts;
  double true_residual_max, bytes;
  unsigned long long digest;
} qwsg_request_result;
int qwsg_server_create(const qwsg_field_spec* fields, int nfields, int links32, int tbc, double csw,
                       const qwsg_solve_params* p, int slots, double mem_budget, qwsg_server** out);
int qwsg_server_run(qwsg_server* srv, const qwsg_request* reqs, int nreq, qwsg_request_result* out);
/* device memory (bytes) the server reserves for a request of nrhs sources on V sites:
 * the pooled slot that serves it (clover inverse, solver vectors for the slot's RHS
 * capacity: 1, 2, 4, 8, 12, 16, 32 or 64) */
int qwsg_server_request_bytes(long long V, int nrhs, double* bytes);
int qwsg_server_destroy(qwsg_server* srv);
int qwsg_synchronize(qwsg_lattice* lat);
/* average device time (ms) of `reps` full FP32/FP16/FP64 dslash launches on the
 * lattice's stream, measured with HIP events */
int qwsg_time_dslash(qwsg_lattice* lat, int prec, int reps, double* ms_per_call);
/* Kernel of the FP32 full operator on 256-site blocks: 0 = neighbour gathers from
 * global memory, 3 or 4 = the block's spinor staged in LDS (compiled for 3 or 4
 * waves per SIMD).  The variants give the same results.  qwsg_autotune_dslash
 * times each over `reps` launches on this lattice, keeps the fastest and returns
 * it in *chosen.  Tuning only; no counterpart in the reference. */
int qwsg_set_dslash_kernel(qwsg_lattice* lat, int kernel);
int qwsg_autotune_dslash(qwsg_lattice* lat, int reps, int* chosen);
/* Block order of the launches over every block (tuning only): 0 index order, 1 t-z
 * pencils inside each XCD's range of workgroups (t and z neighbours close in time, so
 * their faces are read from that XCD's L2).  Results do not depend on it.
 * qwsg_autotune_dslash times both orders with every kernel variant and keeps the
 * fastest pair. */
int qwsg_set_dslash_order(qwsg_lattice* lat, int order);
int qwsg_get_dslash_order(qwsg_lattice* lat, int* order);
/* The solver's SAP settings tuned on this lattice (no counterpart in the reference):
 * one solve of b from x = 0 (x is overwritten) for every (nsap[i], nm[j]) with p's
 * other settings, timed; *best = p with the fastest pair whose FP64 true residual meets
 * p->true_tol (all pairs when true_tol <= 0); seconds[i * nnm + j] per pair (negative:
 * missed the target).  It also keeps the fastest full-operator kernel
 * (qwsg_autotune_dslash). */
int qwsg_autotune_solve(qwsg_lattice* lat, void* x, const void* b, const qwsg_solve_params* p, const int* nsap,
                        int nnsap, const int* nm, int nnm, qwsg_solve_params* best, double* seconds);
/* The SAP block shape tuned for this lattice's extents, FP64 fields and kappa: for
 * each candidate shape (shapes[4i..4i+3], x y z t; nshape of them) a lattice of that
 * shape is built with copies of the fields and solves b (FP64, this lattice's layout)
 * once untimed, then once timed on the host clock with *p.  *best is the index of the
 * fastest; seconds[i] (optional, nshape entries) is negative where the shape does not
 * tile the extents or the timed solve missed p->true_tol.  The lattice itself keeps its
 * blocks: create the production lattice with the chosen shape.  No counterpart in the
 * reference, whose domains are the two x-halves of each process's local lattice
 * (qws.cc:131, 622-631). */
int qwsg_autotune_blocks(qwsg_lattice* lat, const void* b, const qwsg_solve_params* p, const int* shapes, int nshape,
                         int* best, double* seconds);
/* average device time (ms) of one fused SAP half-sweep launch (FP32 or FP16
 * storage), from `reps` times the 2*nsap+2 launches of one preconditioner
 * application on a Gaussian source, measured with HIP events on the lattice's
 * stream.  Measurement only; no counterpart in the reference. */
int qwsg_time_sap(qwsg_lattice* lat, int prec, int reps, int nsap, int nm, double* ms_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* QWS_MI355X_H */
