/*
 * tse_hip.h -- C-ABI of the MI355X (gfx950) TRANSPORTSE SOR / transport core.
 *
 * Plain C: device pointers, sizes, POD structs and int status codes; no
 * torch or HIP types cross this boundary (streams are passed as void*).
 *
 * Device field layout ("checkerboard split"): a field of the nx*ny*nz grid
 * is 2*S doubles, S = nxh*ny*nz, nxh = (nx+1)/2.  Cell (i,j,k) has colour
 * c = (i+j+k) & 1 (0 = red, swept first, tse.py:170-176) and lives at
 *     c*S + (k*ny + j)*nxh + (i >> 1).
 * Each colour half is x-contiguous; slots with 2*(i>>1)+(c^((j+k)&1)) >= nx
 * (odd nx only) are padding and hold 0.  tse_to_split / tse_from_split
 * convert from / to any strided (i,j,k) layout, e.g. numpy C order
 * (nx,ny,nz) with strides (ny*nz, nz, 1) or a species view C[:,:,:,n] with
 * strides (ny*nz*S, nz*S, S).
 *
 * Return codes: 0 = success, < 0 = error (tse_last_error() has the text).
 * Solver non-convergence is not an error: it is reported in tse_result,
 * as the reference returns converged=False (tse.py:742, 916, 1081).
 * One context = one stream; calls on one context must not overlap.
 */
#ifndef TSE_HIP_H
#define TSE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSE_ABI_VERSION 5

enum tse_status {
    TSE_OK = 0,
    TSE_ERR_ARG = -1,     /* invalid argument / shape           */
    TSE_ERR_HIP = -2,     /* HIP runtime error                  */
    TSE_ERR_STATE = -3,   /* grid not set / context misuse      */
    TSE_ERR_NOMEM = -4    /* device allocation failed           */
};

/* Solver forms for tse_compute_p, tse_compute_c and tse_compute_t.
 * FAITHFUL64 is the reference's iteration operation for operation (its
 * update-norm stop test, bit-identical results).  PLAIN64 and MIXED solve
 * the fixed-point system of the same update (the operator of the
 * performance forms below) to the relative residual ||b - A x|| / ||b||
 * <= conv, with FP64 sweeps (PLAIN64) or FP32 sweeps inside FP64 defect
 * correction (MIXED); max_it bounds the sweeps of one attempt, and the
 * time-step control around the attempts is the reference's: accept and
 * grow dt by inc_it (up to max_dt) when an attempt took fewer than max_it/2
 * sweeps, else halve dt and reset the unknown to its entry value, stop with
 * converged = 0 once dt <= min_dt (tse.py:724-740, 896-914, 1065-1079);
 * transport with omega == 0 is accepted after its first sweep launch.  For
 * these forms num_it counts sweeps over all attempts and last_diff is the
 * last relative residual. */
enum tse_p_form {
    TSE_FORM_FAITHFUL64 = 0, /* reference arithmetic, bit-identical      */
    TSE_FORM_PLAIN64 = 1,    /* precomputed operator, FP64 sweeps        */
    TSE_FORM_MIXED = 2       /* FP32 sweeps in FP64 defect correction    */
};

typedef struct tse_ctx tse_ctx;

/* Scalar results of one solver call: the non-array part of
 * (t_P, dt_P, P, Pn, converged, num_it) of compute_P (tse.py:742) and
 * (t, dt, X, converged, num_it) of compute_C / compute_T (tse.py:916,1081). */
typedef struct {
    double t;          /* accepted time t_P / t_C / t_T                 */
    double dt;         /* next step dt_P / dt_C / dt_T                  */
    int32_t converged; /* 1 converged, 0 not (dt fell to min_dt)        */
    int32_t num_it;    /* iterations, including halved attempts         */
    int32_t num_halvings;
    int32_t status;    /* internal final state (1 conv, 3 dt<=min_dt)   */
    double last_diff;  /* last evaluated change norm / residual         */
} tse_result;

/* SOR control parameters: the keyword arguments of compute_P / compute_C /
 * compute_T (tse.py:649-651, 785-787, 921-923). omega is the OMEGA time
 * weight (transport only). */
typedef struct {
    double dt;
    double sor;
    double omega;
    double inc_it;
    double max_dt;
    double min_dt;
    double conv;
    int32_t max_it;
    int32_t form;      /* enum tse_p_form                                */
} tse_sor_params;

/* Inputs of compute_P (tse.py:649). P is updated in place; Pn receives P at
 * entry.  c_r is the frozen module global (tse.py:688).  qx/qy/qz are not
 * used by the reference kernel and are not taken. */
typedef struct {
    double *P, *Pn;
    const double *c_f, *nue, *rho_f, *phi, *Kx, *Ky, *Kz, *BP, *WP, *c_r;
} tse_p_fields;

/* Inputs of compute_C (tse.py:785) for one species; lf_l / lf_t are the
 * frozen dispersion lengths, imblflo is config.C_IMBLFLO (tse.py:865). */
typedef struct {
    double *C, *Cn;
    const double *D_fx, *D_fy, *D_fz, *phi, *rho_f, *qx, *qy, *qz, *BC, *WC,
        *RC, *e, *lf_l, *lf_t;
    int32_t imblflo;
} tse_c_fields;

/* Inputs of compute_T (tse.py:921); l_r, cp_r, rho_r, lt_l, lt_t are the
 * frozen module globals, imblflo is config.T_IMBLFLO (tse.py:1033). */
typedef struct {
    double *T, *Tn;
    const double *l_f, *cp_f, *rho_f, *phi, *qx, *qy, *qz, *BT, *WT, *e,
        *l_r, *cp_r, *rho_r, *lt_l, *lt_t;
    int32_t imblflo;
} tse_t_fields;

/* ---- library / context ------------------------------------------------ */
int tse_abi_version(void);
const char *tse_last_error(void);

/* Create a context on HIP device `device` issuing work on `stream`
 * (a hipStream_t; NULL = the default stream). */
int tse_ctx_create(tse_ctx **out, int device, void *stream);
int tse_ctx_destroy(tse_ctx *ctx);
int tse_ctx_set_stream(tse_ctx *ctx, void *stream);
/* Blocks until all work issued on the context's stream is done. */
int tse_ctx_synchronize(tse_ctx *ctx);

/* Grid geometry (tse.py:82-94): host arrays dx[nx], dy[ny], dz[nz];
 * sin_beta / cos_beta of the slope angle in radians. */
int tse_set_grid(tse_ctx *ctx, int32_t nx, int32_t ny, int32_t nz,
                 const double *dx, const double *dy, const double *dz,
                 double g, double sin_beta, double cos_beta);
/* Doubles per device field (2*S). */
int64_t tse_field_elems(const tse_ctx *ctx);

/* ---- layout conversion ------------------------------------------------ */
/* dst_split[cell] = src[i*si + j*sj + k*sk] (element strides). */
int tse_to_split(tse_ctx *ctx, double *dst_split, const double *src,
                 int64_t si, int64_t sj, int64_t sk);
/* dst[i*si + j*sj + k*sk] = src_split[cell]. */
int tse_from_split(tse_ctx *ctx, double *dst, int64_t si, int64_t sj,
                   int64_t sk, const double *src_split);

/* ---- hot path ------------------------------------------------------------ */
/* compute_P (tse.py:647-742): red-black SOR with the reference's norm
 * schedule and accept / grow / halve / reset control. */
int tse_compute_p(tse_ctx *ctx, const tse_p_fields *f,
                  const tse_sor_params *prm, tse_result *out);

/* compute_P as tse_compute_p (faithful form only), for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * calls no hipMalloc / hipFree: the coefficient records and cell mask live
 * in a fixed device array of 64 MiB that the library holds on each GPU it
 * runs on until the process ends (grids up to about 6.4e5 cells: 104 B of
 * records + 1 mask byte per cell; larger grids return TSE_ERR_ARG).  The
 * context must have been created and its grid set beforehand (those calls
 * allocate).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU, from any context or thread, race and must not be
 * issued.  Results are those of tse_compute_p, bit for bit. */
int tse_compute_p_noalloc(tse_ctx *ctx, const tse_p_fields *f,
                          const tse_sor_params *prm, tse_result *out);

/* compute_q (tse.py:744-780): writes qx, qy, qz in place; the last index in
 * each direction keeps its previous value, as in the reference. */
int tse_compute_q(tse_ctx *ctx, const double *P, const double *BP,
                  const double *Kx, const double *Ky, const double *Kz,
                  double *qx, double *qy, double *qz, const double *nue,
                  const double *rho_f);

/* flow_error (tse.py:409-446): writes e for every cell. */
int tse_flow_error(tse_ctx *ctx, const double *P, const double *Pn,
                   const double *qx, const double *qy, const double *qz,
                   const double *WP, const double *phi, const double *c_f,
                   const double *c_r, const double *rho_f, double dt_P,
                   double *e);

/* compute_C (tse.py:783-916) for one species slice. */
int tse_compute_c(tse_ctx *ctx, const tse_c_fields *f,
                  const tse_sor_params *prm, tse_result *out);

/* compute_T (tse.py:919-1081). */
int tse_compute_t(tse_ctx *ctx, const tse_t_fields *f,
                  const tse_sor_params *prm, tse_result *out);

/* smolarkiewicz (tse.py:1083-1116): anti-diffusive velocities of the
 * Smolarkiewicz correction from the field C (a species slice or T) and the
 * fluxes; vx, vy, vz are updated in place (the last index of an axis keeps
 * its previous value, as in the reference). */
int tse_smolarkiewicz(tse_ctx *ctx, const double *C, const double *qx,
                      const double *qy, const double *qz, double *vx,
                      double *vy, double *vz, double dt, double eps, double Sc);

/* l2_norm (tse.py:273-281) of two device fields: a parallel FP64
 * reduction (rounding differs from Numba's sequential sum by O(1e-16)). */
int tse_l2_norm(tse_ctx *ctx, const double *New, const double *Old,
                double *out);
/* The same reduction, enqueued without waiting: the result is read by
 * tse_reduction_result(ctx, out, 1).  No counterpart in the reference; for
 * callers that drive many contexts on one stream and wait once for all. */
int tse_l2_norm_async(tse_ctx *ctx, const double *New, const double *Old);
/* Waits for the context's stream and returns the last enqueued reduction
 * (tse_l2_norm_async: n = 1; tse_mass_sums_async: n = 3). */
int tse_reduction_result(tse_ctx *ctx, double *out, int32_t n);

/* l2_norm on host arrays of n doubles with Numba's sequential C-order sums
 * (used by the SIA density test, tse.py:1426). Returns the norm, or 1.0 and
 * *is_nan = 1 on NaN. */
double tse_host_l2_norm_seq(const double *New, const double *Old, int64_t n,
                            int32_t *is_nan);

/* ---- performance forms of the pressure solve ------------------------------ */
/* The compute_P system (tse.py:657-697 at its fixed point) scaled by the cell
 * volume is symmetric: each cell stores its three "+" face transmissibilities,
 * 1/diagonal and the right-hand side (natural x-contiguous layout, owned by
 * the context; its rows are nx rounded up to even: an odd nx gets one
 * inactive column).  Sweeps are fused red+black z-marches (NS sweeps per launch);
 * convergence is the relative residual ||b - A p|| / ||b|| in FP64.
 *   TSE_FORM_PLAIN64: FP64 sweeps on p.
 *   TSE_FORM_MIXED:   FP32 sweeps on the correction A e = r, r = b - A p in
 *                     FP64, p += e in FP64 (defect correction). */
/* Launch plan of the fused sweep kernel: ns sweeps per launch, tile of
 * 64 x by columns with ry tile rows per thread (blocks of 64 x by/ry
 * threads), zchunk z planes per block (<= 0: whole column), variant 0
 * (x neighbours through LDS), 1 (x neighbours exchanged across the lanes
 * of a wavefront; smaller LDS ring), 2 (as 1, with the register rings
 * rotated by unrolling and buffer loads; fields of at most 1 GiB, else
 * TSE_ERR_ARG), 3 (FP32 only, as 2 with an x-pair of cells per thread, so
 * that every lane updates in every phase; pressure rows are even), 5 or 7
 * (below).  Supported (ns, by, ry):
 *   variant 0: (1,8,1) (1,16,1) (2,16,1) (1,24,2); FP32 sweeps also
 *              (1,32,2) (2,32,2) (2,32,4) (3,16,1)
 *   variant 1: (1,16,1) (2,16,1) (1,24,2); FP32 sweeps also (2,24,2)
 *              (2,32,2) (3,16,1) (3,24,2)
 *   variant 2: as variant 1, FP32 sweeps also (3,32,2)
 *   variant 3: (2,32,2) (3,32,2)
 *   variant 5: (2,32,2) (3,32,2), FP32 pressure sweeps only (single
 *              systems and batches): variant 3's cell pairs on an inner
 *              operator stored as bfloat16 (the three faces and the
 *              storage term of each cell; the diagonal is summed in the
 *              kernel from the rounded values), a graph Laplacian within a
 *              factor 1 +- 2^-8 of the FP64 one, so the defect correction
 *              keeps its stop test and sweep count; 20 instead of 28 bytes
 *              per cell per launch; at most 2^27 cells
 *   variant 7: (2,32,2) (3,32,2), single FP32 pressure systems only (a
 *              batch rejects it): variant 5 with the correction's rhs
 *              stored as FP16, scaled by a power of two near 1 / RMS of
 *              the residual, rounded from its FP32 value with a dither
 *              fixed by the cell index and clamped to +-65504; 18 bytes
 *              per cell per launch; at and above the optimal factor the
 *              same sweeps as variant 5, below it 1.5-6 % more
 * Transport systems: variant 0 or 2, (1,8,1) (1,16,1) (2,16,1), FP32
 * (3,16,1). */
typedef struct {
    int32_t ns, by, ry, zchunk, variant;
} tse_sweep_plan;

typedef struct {
    int32_t form;       /* TSE_FORM_PLAIN64 or TSE_FORM_MIXED              */
    tse_sweep_plan plan;
    int32_t inner;      /* sweeps per residual check / correction cycle     */
    int32_t max_sweeps; /* stop after this many sweeps                      */
    int32_t cycle_mode; /* TSE_CYCLE_FIXED or TSE_CYCLE_ADAPTIVE (solve only) */
    int32_t adapt_sor;  /* 1 (single solves only; a batch rejects it): after
                         * each cycle from the second on whose contraction
                         * shows sor below the optimum, continue at the
                         * estimate of tse_sor_model (rounded up to 1/128 in
                         * log2(2 - sor); at most twice per solve); 0: sor
                         * throughout                                        */
    double sor;         /* over-relaxation factor (the starting one if
                         * adapt_sor)                                        */
    double tol;         /* relative residual target                         */
} tse_fast_params;
/* Cycle lengths of tse_fast_*_solve.  FIXED: every cycle is `inner` sweeps.
 * ADAPTIVE: the first cycle is `inner` sweeps; after each residual check
 * the rate of the last cycle predicts the sweeps still needed to reach tol,
 * and the next cycle is that many (plus 2 %), rounded up to a multiple of
 * the plan's ns and kept within [inner, 8*inner]; the stop test is the same
 * (relative residual <= tol), with fewer FP64 residual passes.  The batched
 * solves use fixed cycles. */
enum { TSE_CYCLE_FIXED = 0, TSE_CYCLE_ADAPTIVE = 1 };

typedef struct {
    int32_t sweeps;     /* sweeps performed                                 */
    int32_t cycles;     /* residual checks / correction cycles              */
    int32_t converged;  /* 1 if rel_residual <= tol                          */
    int32_t pad;
    double initial_rel; /* relative residual of the initial iterate         */
    double rel_residual;
    double contraction; /* residual contraction per sweep over the last
                         * cycle after the first, if that cycle of n sweeps
                         * at factor sor spans n |log(sor - 1)| >= 1/4
                         * (shorter ones are dominated by their restart
                         * transient); else 0                               */
    double sor;         /* over-relaxation factor of the last cycle (single
                         * solves; the batch's or the member's in a status) */
} tse_fast_result;

/* Young's relation for red/black SOR on a consistently ordered matrix (the
 * 7-point pressure operator in red/black order): from the residual
 * contraction per sweep `contraction` observed at factor `omega`, returns 1
 * and the optimal factor in *omega_opt when the contraction is clearly
 * slower than omega - 1 (omega below the optimum; the estimate is capped at
 * 2 - 2^-10 when it cannot be resolved), 0 when it is about omega - 1 (at or
 * above the optimum), TSE_ERR_ARG for inputs outside (0, 1) and (1, 2).
 * Host only.  No counterpart in the reference: the adaptive-SOR estimate of
 * the performance forms (Young 1971, Hageman and Young 1981). */
int tse_sor_model(double omega, double contraction, double *omega_opt);

/* Assembles the operator from compute_P's inputs at time step dt (Pn = the
 * storage-term reference pressure; NULL uses P) and loads P as the iterate. */
int tse_fast_p_setup(tse_ctx *ctx, const tse_p_fields *f, double dt);
/* Solves to prm->tol (or prm->max_sweeps) from the current iterate; a
 * cycle that would run past max_sweeps is shortened to end there (rounded
 * up to the plan's ns).  Batched members instead stop at the first cycle end
 * at or beyond max_sweeps. */
int tse_fast_p_solve(tse_ctx *ctx, const tse_fast_params *prm,
                     tse_fast_result *out);
/* Relative residual of the current iterate. */
int tse_fast_p_residual(tse_ctx *ctx, double *rel);
/* Copies the assembled FP64 operator (natural layout with rows of nx
 * rounded up to even, nxf * ny * nz doubles each: the "+" faces Wx, Wy, Wz,
 * the diagonal D and the rhs b) into device buffers;
 * a NULL destination is skipped.  No counterpart in the reference: an
 * inspection hook for tests and tools. */
int tse_fast_p_operator(tse_ctx *ctx, double *Wx, double *Wy, double *Wz,
                        double *D, double *b);
/* Writes the current iterate into a split-layout field. */
int tse_fast_p_result(tse_ctx *ctx, double *P);
/* Timing hook: n_sweeps fused sweeps (FP64 on p, or fp32 != 0: FP32 on the
 * correction with the last residual), no residual checks. */
int tse_fast_p_sweeps(tse_ctx *ctx, const tse_sweep_plan *plan,
                      int32_t n_sweeps, double sor, int32_t fp32);

/* ---- performance forms of the transport solves ---------------------------- */
/* compute_C (tse.py:783-916) and compute_T (tse.py:919-1081) at their fixed
 * point: (R - omega M) x - omega sum_f a_f x_f = X + R Xn + W - E with the
 * reference's coefficients (the faithful records, evaluated with the
 * reference arithmetic), non-symmetric (upwind advection).  Each cell keeps
 * its six neighbour coefficients, 1/diagonal and rhs in the natural layout;
 * the sweeps are fused red/black z-marches as for pressure, the stop test is
 * the relative residual, forms PLAIN64 or MIXED.  Supported plans (ns, by,
 * ry): (1,8,1) (1,16,1) (2,16,1), and for FP32 also (3,16,1).  One transport
 * system per context: a setup replaces the previous one.
 * Setup copies the unknown at entry (C or T) into Cn / Tn, assembles at dt
 * with time weighting omega, and loads the unknown as the iterate. */
int tse_fast_c_setup(tse_ctx *ctx, const tse_c_fields *f, double dt,
                     double omega);
int tse_fast_t_setup(tse_ctx *ctx, const tse_t_fields *f, double dt,
                     double omega);
int tse_fast_tr_solve(tse_ctx *ctx, const tse_fast_params *prm,
                      tse_fast_result *out);
int tse_fast_tr_residual(tse_ctx *ctx, double *rel);
int tse_fast_tr_result(tse_ctx *ctx, double *X);
int tse_fast_tr_sweeps(tse_ctx *ctx, const tse_sweep_plan *plan,
                       int32_t n_sweeps, double sor, int32_t fp32);

/* ---- batched performance-form solves ------------------------------------- */
/* No counterpart in the reference (its runs are one grid per process): the
 * ensemble and serving workloads solve many independent grids at once.  A
 * batch holds member contexts whose fast system of one kind (0 = pressure,
 * set up by tse_fast_p_setup; 1 = transport, by tse_fast_c/t_setup) is ready.
 * A cycle runs, for every active member, the same `inner` sweeps, residual
 * and stop test as tse_fast_p_solve / tse_fast_tr_solve run per cycle, so a
 * member's result does not depend on the other members.  The sweeps of all
 * members run in one persistent launch per sweep group that takes (member,
 * tile) work items from an atomic queue.  Members' solves stop at their own
 * tolerance; a stopped member is skipped until it is re-armed or removed.
 * Calls are not thread-safe per batch; a member context must not be used by
 * other fast-form calls of the same kind while it is in a batch (use
 * tse_batch_writeback first to read its result). */
typedef struct tse_batch tse_batch;
int tse_batch_create(tse_batch **out, int32_t device, void *stream);
int tse_batch_destroy(tse_batch *b);
/* Sets kind, form, plan, inner, max_sweeps, sor, tol; the batch must be empty. */
int tse_batch_begin(tse_batch *b, int32_t kind, const tse_fast_params *prm);
/* Adds a member and evaluates its initial residual; returns its slot (>= 0). */
int tse_batch_add(tse_batch *b, tse_ctx *member);
/* Restarts a member after its fast system was set up again (next solve). */
int tse_batch_rearm(tse_batch *b, int32_t slot);
/* Restarts n members at once (one synchronisation). */
int tse_batch_rearm_many(tse_batch *b, const int32_t *slots, int32_t n);
/* Hands the member's iterate back to its context (tse_fast_*_result). */
int tse_batch_writeback(tse_batch *b, int32_t slot);
/* Over-relaxation factor of one member (default: the batch's sor); takes
   effect at the next cycle or re-arm, which uploads it without a
   synchronisation of its own. */
int tse_batch_set_sor(tse_batch *b, int32_t slot, double sor);
/* Writes back and removes a member. */
int tse_batch_remove(tse_batch *b, int32_t slot);
/* One cycle for all active members; *n_active = members still active.
 * use_graph != 0 replays a hipGraph of the cycle, captured on first use and
 * kept while the member set is unchanged (the stream must not be the null
 * stream). */
int tse_batch_cycle(tse_batch *b, int32_t use_graph, int32_t *n_active);
/* Cycles until no member is active; *cycles = cycles run. */
int tse_batch_run(tse_batch *b, int32_t use_graph, int32_t *cycles);
/* Member state: sweeps, cycles, converged, pad = still active, initial and
 * last relative residual. */
int tse_batch_status(tse_batch *b, int32_t slot, tse_fast_result *out);

/* ---- device-resident SIA step pieces ------------------------------------ */
/* Closed-form linear fluid properties of the Picard update (the
 * fluid_properties callback of tse.py:178-229, 1422-1423, for configs whose
 * EOS is linear): rho_f = rho0 + drho_dc * C and, if l_f is not NULL,
 * l_f = lf0 * (1 + dlf_dt * (T - t_ref)).  C is the salt species field. */
typedef struct {
    double rho0, drho_dc, lf0, dlf_dt, t_ref;
} tse_eos_params;
int tse_linear_eos(tse_ctx *ctx, const double *C, const double *T,
                   double *rho_f, double *l_f, const tse_eos_params *p);
/* FP64 mass-balance sums (the mass_balance callback, tse.py:1461-1463, of
 * the benchmark configs): out3 = [sum C phi V, sum rho_f phi V, outflow
 * rate of C through the two x faces (upwind, face area dy*dz)]. */
int tse_mass_sums(tse_ctx *ctx, const double *C, const double *rho_f,
                  const double *phi, const double *qx, double *out3);
/* The same sums enqueued without waiting (tse_reduction_result, n = 3). */
int tse_mass_sums_async(tse_ctx *ctx, const double *C, const double *rho_f,
                        const double *phi, const double *qx);
/* peclet_C (tse.py:333-360): for cells with BC == 0,
 * PeC[f*3 + a] = |q_a| / max_s D_a[f*nspec + s] * d_a (NaN where that
 * maximum is 0); other cells keep their PeC.  PeC and D_fx/y/z are device
 * arrays in the host layout (C order (nx, ny, nz, 3) and (nx, ny, nz,
 * nspec)), q and BC split-layout fields.  out3 = max over all cells of
 * PeC[..., a], NaN if any element is NaN (numpy.max). */
int tse_peclet_c(tse_ctx *ctx, const double *qx, const double *qy,
                 const double *qz, const double *D_fx, const double *D_fy,
                 const double *D_fz, int32_t nspec, const double *BC,
                 double *PeC, double *out3);
/* peclet_T (tse.py:364-383): for cells with BT == 0,
 * PeT[f*3 + a] = |q_a| rho_f cp_f / ((1-phi) l_r + phi l_f) * d_a; other
 * cells keep their PeT.  Fields split-layout, PeT host layout; out3 as for
 * tse_peclet_c. */
int tse_peclet_t(tse_ctx *ctx, const double *qx, const double *qy,
                 const double *qz, const double *rho_f, const double *cp_f,
                 const double *phi, const double *l_r, const double *l_f,
                 const double *BT, double *PeT, double *out3);
/* The three maxima of the SIA driver's CFL limit (tse.py:1241-1246),
 * dt_CFL = 0.9 / (out3[0] + out3[1] + out3[2]): max |qx|/phi/dx,
 * |qy|/phi/dy, |qz|/phi/dz with the reference's reshape of the first two
 * arrays to (ny, nz, nx) and (nx, nz, ny) before the division, i.e. C-order
 * element f divided by dx[f % nx] and dy[f % ny].  NaN-propagating. */
int tse_cfl_maxima(tse_ctx *ctx, const double *qx, const double *qy,
                   const double *qz, const double *phi, double *out3);

/* ---- graphs of a context's device work ------------------------------------ */
/* Captures the work that calls issue on the context's stream between
 * tse_graph_begin and tse_graph_end (no call that synchronises may come in
 * between) and replays it with tse_graph_launch.  Host-side effects of the
 * captured calls happen at capture time only: after replaying a fast-form
 * setup, call tse_fast_mark_loaded to restore what the setup sets on the
 * host (the iterate is in the first buffer and ready), and a captured
 * tse_fast_*_result reads the iterate buffer that was current at capture,
 * so key such graphs by tse_fast_current.  No counterpart in the reference
 * (its step is Python calls); these serve callers that replay a time step's
 * device work, like the batched ensemble. */
typedef struct tse_graph tse_graph;
int tse_graph_begin(tse_ctx *ctx);
int tse_graph_end(tse_ctx *ctx, tse_graph **out);
int tse_graph_launch(tse_graph *g, tse_ctx *ctx);
int tse_graph_destroy(tse_graph *g);
/* Host-side state of a fast system (kind 0 pressure, 1 transport) after its
 * setup: iterate in buffer 0, ready. */
int tse_fast_mark_loaded(tse_ctx *ctx, int32_t kind);
/* Which of the two iterate buffers (0 or 1) holds the current iterate. */
int tse_fast_current(tse_ctx *ctx, int32_t kind, int32_t *cur);

/* ---- measurement hooks -------------------------------------------------- */
/* Runs n_iter full red+black iterations of the faithful compute_P sweep at
 * fixed dt (no norm, no control), for timing and sweep-level parity.
 * assemble != 0 first evaluates the coefficient records from f (as
 * tse_compute_p does on entry); assemble == 0 reuses the records of the
 * previous call on this context (same fields and dt). */
int tse_p_sweeps(tse_ctx *ctx, const tse_p_fields *f, double dt, double sor,
                 int32_t n_iter, int32_t assemble);

#ifdef __cplusplus
}
#endif
#endif
