/*
 * tse_oracle.h -- CPU restatement of TRANSPORTSE's SOR / transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests compare the HIP path with this
 * library, and it is the timed CPU baseline; nothing in the product (transportse_amd/)
 * may link, load or call it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * Arrays use the reference layout: numpy C order of shape (nx, ny, nz), so
 * k (z) is the contiguous axis and cell (i,j,k) lives at (i*ny + j)*nz + k.
 * All arithmetic is IEEE binary64 with the operand order of tse.py and no
 * FMA contraction (built with -ffp-contract=off), which is what Numba
 * produces with fastmath=False.
 */
#ifndef TSE_ORACLE_H
#define TSE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid geometry and the module globals Numba freezes at first compilation
 * (tse.py:82-94, 105-118). sin_beta / cos_beta are computed by the caller
 * from numpy.radians(config.beta) so every backend shares the same bits. */
typedef struct {
    int nx, ny, nz;
    const double *dx, *dy, *dz;   /* lengths nx, ny, nz */
    double g;
    double sin_beta, cos_beta;
} tse_or_grid;

/* Result of an SOR solver call, the scalar part of the reference's return
 * tuples (tse.py:742, 916, 1081). */
typedef struct {
    double t;          /* accepted time t_P / t_C / t_T            */
    double dt;         /* next time step dt_P / dt_C / dt_T        */
    int converged;     /* 1 if converged                           */
    int num_it;        /* total iterations incl. halved attempts   */
    int num_halvings;  /* number of dt halvings (print lines)      */
    double last_diff;  /* last l2 change norm evaluated            */
} tse_or_result;

/* ---- helpers (tse.py:248-310) ------------------------------------------ */
double tse_or_harm(const tse_or_grid *g, const double *H, int i1, int j1,
                   int k1, int i2, int j2, int k2, int n);
double tse_or_amean(const tse_or_grid *g, const double *X, int i1, int j1,
                    int k1, int i2, int j2, int k2, int n);
double tse_or_alpha(const tse_or_grid *g, int i1, int j1, int k1, int i2,
                    int j2, int k2, const double *nue, const double *rho_f,
                    const double *K, int d1, int d2, int n);
double tse_or_beta_g(const tse_or_grid *g, int i1, int j1, int k1, int i2,
                     int j2, int k2, const double *rho_f, int d);
/* Returns the norm; *is_nan is set (and 1.0 returned) on NaN. */
double tse_or_l2_norm(const double *New, const double *Old, size_t n,
                      int *is_nan);
/* D_xl, D_yl, D_zl, D_xr, D_yr, D_zr (tse.py:460-644). */
void tse_or_dispersivity(const tse_or_grid *g, const double *qx,
                         const double *qy, const double *qz,
                         const double *l_l, const double *l_t,
                         const double *B, int i, int j, int k, double out[6]);

/* ---- solvers ----------------------------------------------------------- */
/* Progress report of long compute_P calls: cb(num_it) every `every`
 * iterations (NULL or every <= 0: off).  Changes no value. */
void tse_or_set_progress(void (*cb)(int), int every);

/* compute_P (tse.py:647-742). P is updated in place (it ends holding the
 * returned P), Pn receives the copy of P at entry. c_r is the frozen
 * module global. */
int tse_or_compute_p(const tse_or_grid *g, double *P, double *Pn,
                     const double *c_f, const double *nue,
                     const double *rho_f, const double *phi,
                     const double *Kx, const double *Ky, const double *Kz,
                     const double *BP, const double *WP, const double *c_r,
                     double dt_P, double SOR, int max_it, double inc_it,
                     double max_dt, double min_dt, double P_conv,
                     tse_or_result *out);

/* One red (colour 0) or black (colour 1) half-sweep of compute_P's solve()
 * with fixed dt_P (tse.py:705-716). */
void tse_or_p_half_sweep(const tse_or_grid *g, double *P, const double *Pn,
                         const double *c_f, const double *nue,
                         const double *rho_f, const double *phi,
                         const double *Kx, const double *Ky, const double *Kz,
                         const double *BP, const double *WP,
                         const double *c_r, double dt_P, double SOR,
                         int colour);

/* compute_q (tse.py:744-780): writes qx, qy, qz in place. */
void tse_or_compute_q(const tse_or_grid *g, const double *P, const double *BP,
                      const double *Kx, const double *Ky, const double *Kz,
                      double *qx, double *qy, double *qz, const double *nue,
                      const double *rho_f);

/* flow_error (tse.py:409-446): writes a fresh e (all cells). */
void tse_or_flow_error(const tse_or_grid *g, const double *P,
                       const double *Pn, const double *qx, const double *qy,
                       const double *qz, const double *WP, const double *phi,
                       const double *c_f, const double *c_r,
                       const double *rho_f, double dt_P, double *e);

/* compute_C (tse.py:783-916) for one species slice (contiguous copy).
 * lf_l / lf_t are frozen globals, imblflo is config.C_IMBLFLO. */
int tse_or_compute_c(const tse_or_grid *g, double *C, double *Cn_out,
                     const double *D_fx, const double *D_fy,
                     const double *D_fz, const double *phi,
                     const double *rho_f, const double *qx, const double *qy,
                     const double *qz, const double *BC, const double *WC,
                     const double *RC, const double *e, const double *lf_l,
                     const double *lf_t, int imblflo, double dt_C,
                     double omega, double SOR, int max_it, double inc_it,
                     double max_dt, double min_dt, double C_conv,
                     tse_or_result *out);

/* compute_T (tse.py:919-1081). l_r, cp_r, rho_r, lt_l, lt_t are frozen
 * globals, imblflo is config.T_IMBLFLO. */
int tse_or_compute_t(const tse_or_grid *g, double *T, double *Tn_out,
                     const double *l_f, const double *cp_f,
                     const double *rho_f, const double *phi,
                     const double *qx, const double *qy, const double *qz,
                     const double *BT, const double *WT, const double *e,
                     const double *l_r, const double *cp_r,
                     const double *rho_r, const double *lt_l,
                     const double *lt_t, int imblflo, double dt_T,
                     double omega, double SOR, int max_it, double inc_it,
                     double max_dt, double min_dt, double T_conv,
                     tse_or_result *out);

/* Number of sweep threads (OpenMP); 0 leaves the OpenMP default. */
/* smolarkiewicz (tse.py:1083-1116); vx, vy, vz updated in place. */
void tse_or_smolarkiewicz(const tse_or_grid *g, const double *C,
                          const double *qx, const double *qy, const double *qz,
                          double *vx, double *vy, double *vz, double dt,
                          double eps, double Sc);
void tse_or_set_threads(int n);
int tse_or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
