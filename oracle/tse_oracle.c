/*
 * tse_oracle.c -- CPU restatement of TRANSPORTSE's SOR / transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY (what the parity tests compare with, and the
 * timed CPU baseline; the product never links it).
 *
 * Each function restates the Numba kernel of tse.py cited above it, with the
 * same operand order, the same branch structure and the same control flow.
 * Built with -O3 -ffp-contract=off and without -ffast-math, so every
 * arithmetic operation is one correctly rounded IEEE binary64 operation, as
 * in Numba with fastmath=False.
 *
 * Parallelism: a colour half-sweep writes only cells of that colour and
 * reads only the other colour plus the cell itself (tse.py:170-176), so the
 * cell order inside a half-sweep cannot change any result; OpenMP splits the
 * i axis. Reductions (l2_norm) are sequential in C order, matching Numba's
 * nditer loop for numpy.sum inside a non-parallel njit function.
 */
#include "tse_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(g, i, j, k) \
    (((size_t)(i) * (size_t)(g)->ny + (size_t)(j)) * (size_t)(g)->nz + (size_t)(k))

/* Python's min(q, 0) / max(q, 0): the first argument wins unless the second
 * compares strictly smaller / larger (tse.py:812-813 etc.). */
static inline double py_min0(double q) { return (0.0 < q) ? 0.0 : q; }
static inline double py_max0(double q) { return (0.0 > q) ? 0.0 : q; }

static inline double py_min(double a, double b) { return (b < a) ? b : a; }

static int g_threads = 0;

void tse_or_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int tse_or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- helpers ------------------------------------------------------------ */

/* harm (tse.py:297-310): distance-weighted harmonic mean, 0 when either
 * operand is 0. */
double tse_or_harm(const tse_or_grid *g, const double *H, int i1, int j1,
                   int k1, int i2, int j2, int k2, int n) {
    double h1 = H[IDX(g, i1, j1, k1)], h2 = H[IDX(g, i2, j2, k2)];
    if (h1 != 0 && h2 != 0) {
        if (n == 0)
            return (g->dx[i1] + g->dx[i2]) / (g->dx[i1] / h1 + g->dx[i2] / h2);
        else if (n == 1)
            return (g->dy[j1] + g->dy[j2]) / (g->dy[j1] / h1 + g->dy[j2] / h2);
        else
            return (g->dz[k1] + g->dz[k2]) / (g->dz[k1] / h1 + g->dz[k2] / h2);
    }
    return 0.0;
}

/* amean (tse.py:284-294): distance-weighted arithmetic mean. */
double tse_or_amean(const tse_or_grid *g, const double *X, int i1, int j1,
                    int k1, int i2, int j2, int k2, int n) {
    double x1 = X[IDX(g, i1, j1, k1)], x2 = X[IDX(g, i2, j2, k2)];
    if (n == 0)
        return (x1 * g->dx[i1] + x2 * g->dx[i2]) / (g->dx[i1] + g->dx[i2]);
    else if (n == 1)
        return (x1 * g->dy[j1] + x2 * g->dy[j2]) / (g->dy[j1] + g->dy[j2]);
    else
        return (x1 * g->dz[k1] + x2 * g->dz[k2]) / (g->dz[k1] + g->dz[k2]);
}

static inline const double *spacing(const tse_or_grid *g, int n) {
    return n == 0 ? g->dx : (n == 1 ? g->dy : g->dz);
}

/* alpha (tse.py:249-260): harm(K)/harm(nue)/(d[d2]*(d[d1]+d[d2])*0.5)
 * *harm(rho_f); d2 is the centre cell. */
double tse_or_alpha(const tse_or_grid *g, int i1, int j1, int k1, int i2,
                    int j2, int k2, const double *nue, const double *rho_f,
                    const double *K, int d1, int d2, int n) {
    const double *d = spacing(g, n);
    double hk = tse_or_harm(g, K, i1, j1, k1, i2, j2, k2, n);
    double hn = tse_or_harm(g, nue, i1, j1, k1, i2, j2, k2, n);
    double hr = tse_or_harm(g, rho_f, i1, j1, k1, i2, j2, k2, n);
    return hk / hn / (d[d2] * (d[d1] + d[d2]) * 0.5) * hr;
}

/* beta_g (tse.py:263-271): gravity head on an x (d == 0) or z face. */
double tse_or_beta_g(const tse_or_grid *g, int i1, int j1, int k1, int i2,
                     int j2, int k2, const double *rho_f, int d) {
    if (d == 0)
        return tse_or_amean(g, rho_f, i1, j1, k1, i2, j2, k2, 0) * g->g * 0.5 *
               (g->dx[i1] + g->dx[i2]) * g->sin_beta;
    return tse_or_amean(g, rho_f, i1, j1, k1, i2, j2, k2, 2) * g->g * 0.5 *
           (g->dz[k1] + g->dz[k2]) * g->cos_beta;
}

/* l2_norm (tse.py:273-281): sqrt(sum((New-Old)**2))/sqrt(sum(Old**2)+1e-200)
 * with both sums accumulated sequentially in C order. */
double tse_or_l2_norm(const double *New, const double *Old, size_t n,
                      int *is_nan) {
    double s1 = 0.0, s2 = 0.0;
    for (size_t p = 0; p < n; ++p) {
        double d = New[p] - Old[p];
        s1 += d * d;
    }
    for (size_t p = 0; p < n; ++p) s2 += Old[p] * Old[p];
    double l2 = sqrt(s1) / sqrt(s2 + 1e-200);
    if (isnan(l2)) {
        if (is_nan) *is_nan = 1;
        return 1.0;
    }
    if (is_nan) *is_nan = 0;
    return l2;
}

/* One transversal average (tse.py:472-481 and its 11 siblings):
 * (q1 + q2 + q3 + q4) / c with q3 = q4 = 0 and c = 2 at the low edge. */
static inline double tavg(double q1, double q2, int edge_ok, double q3,
                          double q4) {
    if (edge_ok) return (q1 + q2 + q3 + q4) / 4.0;
    return (q1 + q2 + 0.0 + 0.0) / 2.0;
}

/* dispersivity (tse.py:460-644). */
void tse_or_dispersivity(const tse_or_grid *g, const double *qx,
                         const double *qy, const double *qz,
                         const double *l_l, const double *l_t,
                         const double *B, int i, int j, int k, double out[6]) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double qxl = 0, qyl = 0, qzl = 0, qxr = 0, qyr = 0, qzr = 0;
    double qxyl = 0, qxzl = 0, qyxl = 0, qyzl = 0, qzxl = 0, qzyl = 0;
    double qxyr = 0, qxzr = 0, qyxr = 0, qyzr = 0, qzxr = 0, qzyr = 0;
#define Q(a, ii, jj, kk) (a)[IDX(g, ii, jj, kk)]
    if (i != 0) { /* left x face */
        qxl = Q(qx, i - 1, j, k);
        if (!(Q(B, i - 1, j, k) != 0)) {
            qxyl = tavg(Q(qy, i - 1, j, k), Q(qy, i, j, k), j != 0,
                        j != 0 ? Q(qy, i, j - 1, k) : 0.0,
                        j != 0 ? Q(qy, i - 1, j - 1, k) : 0.0);
            qxzl = tavg(Q(qz, i - 1, j, k), Q(qz, i, j, k), k != 0,
                        k != 0 ? Q(qz, i, j, k - 1) : 0.0,
                        k != 0 ? Q(qz, i - 1, j, k - 1) : 0.0);
        }
    }
    if (i != nx - 1) { /* right x face */
        qxr = Q(qx, i, j, k);
        if (!(Q(B, i + 1, j, k) != 0)) {
            qxyr = tavg(Q(qy, i, j, k), Q(qy, i + 1, j, k), j != 0,
                        j != 0 ? Q(qy, i + 1, j - 1, k) : 0.0,
                        j != 0 ? Q(qy, i, j - 1, k) : 0.0);
            qxzr = tavg(Q(qz, i, j, k), Q(qz, i + 1, j, k), k != 0,
                        k != 0 ? Q(qz, i + 1, j, k - 1) : 0.0,
                        k != 0 ? Q(qz, i, j, k - 1) : 0.0);
        }
    }
    if (j != 0) { /* left y face */
        qyl = Q(qy, i, j - 1, k);
        if (!(Q(B, i, j - 1, k) != 0)) {
            qyxl = tavg(Q(qx, i, j, k), Q(qx, i, j - 1, k), i != 0,
                        i != 0 ? Q(qx, i - 1, j - 1, k) : 0.0,
                        i != 0 ? Q(qx, i - 1, j, k) : 0.0);
            qyzl = tavg(Q(qz, i, j - 1, k), Q(qz, i, j, k), k != 0,
                        k != 0 ? Q(qz, i, j, k - 1) : 0.0,
                        k != 0 ? Q(qz, i, j - 1, k - 1) : 0.0);
        }
    }
    if (j != ny - 1) { /* right y face */
        qyr = Q(qy, i, j, k);
        if (!(Q(B, i, j + 1, k) != 0)) {
            qyxr = tavg(Q(qx, i, j + 1, k), Q(qx, i, j, k), i != 0,
                        i != 0 ? Q(qx, i - 1, j, k) : 0.0,
                        i != 0 ? Q(qx, i - 1, j + 1, k) : 0.0);
            qyzr = tavg(Q(qz, i, j, k), Q(qz, i, j + 1, k), k != 0,
                        k != 0 ? Q(qz, i, j + 1, k - 1) : 0.0,
                        k != 0 ? Q(qz, i, j, k - 1) : 0.0);
        }
    }
    if (k != 0) { /* bottom z face */
        qzl = Q(qz, i, j, k - 1);
        if (!(Q(B, i, j, k - 1) != 0)) {
            qzxl = tavg(Q(qx, i, j, k), Q(qx, i, j, k - 1), i != 0,
                        i != 0 ? Q(qx, i - 1, j, k - 1) : 0.0,
                        i != 0 ? Q(qx, i - 1, j, k) : 0.0);
            qzyl = tavg(Q(qy, i, j, k), Q(qy, i, j, k - 1), j != 0,
                        j != 0 ? Q(qy, i, j - 1, k - 1) : 0.0,
                        j != 0 ? Q(qy, i, j - 1, k) : 0.0);
        }
    }
    if (k != nz - 1) { /* top z face */
        qzr = Q(qz, i, j, k);
        if (!(Q(B, i, j, k + 1) != 0)) {
            qzxr = tavg(Q(qx, i, j, k + 1), Q(qx, i, j, k), i != 0,
                        i != 0 ? Q(qx, i - 1, j, k) : 0.0,
                        i != 0 ? Q(qx, i - 1, j, k + 1) : 0.0);
            qzyr = tavg(Q(qy, i, j, k + 1), Q(qy, i, j, k), j != 0,
                        j != 0 ? Q(qy, i, j - 1, k) : 0.0,
                        j != 0 ? Q(qy, i, j - 1, k + 1) : 0.0);
        }
    }
    const double ll = Q(l_l, i, j, k), lt = Q(l_t, i, j, k);
#undef Q
    out[0] = ll * fabs(qxl) + lt * (fabs(qxyl) + fabs(qxzl)); /* D_xl */
    out[1] = ll * fabs(qyl) + lt * (fabs(qyxl) + fabs(qyzl)); /* D_yl */
    out[2] = ll * fabs(qzl) + lt * (fabs(qzxl) + fabs(qzyl)); /* D_zl */
    out[3] = ll * fabs(qxr) + lt * (fabs(qxyr) + fabs(qxzr)); /* D_xr */
    out[4] = ll * fabs(qyr) + lt * (fabs(qyxr) + fabs(qyzr)); /* D_yr */
    out[5] = ll * fabs(qzr) + lt * (fabs(qzxr) + fabs(qzyr)); /* D_zr */
}

/* ---- compute_P ---------------------------------------------------------- */

/* solve() of compute_P (tse.py:657-697) for one cell. */
static inline double p_solve(const tse_or_grid *g, const double *P,
                             const double *Pn, const double *c_f,
                             const double *nue, const double *rho_f,
                             const double *phi, const double *Kx,
                             const double *Ky, const double *Kz,
                             const double *WP, const double *c_r, double dt_P,
                             double SOR, int i, int j, int k) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double A_ = 0, B_ = 0, C_ = 0, D_ = 0, E_ = 0, G_ = 0, I_ = 0, J_ = 0,
           K_ = 0, L_ = 0, Anew = 0, Cnew = 0, Enew = 0, Gnew = 0, Inew = 0,
           Knew = 0;
    if (i != nx - 1) {
        A_ = tse_or_alpha(g, i + 1, j, k, i, j, k, nue, rho_f, Kx, i + 1, i, 0);
        Anew = A_ * P[IDX(g, i + 1, j, k)];
        B_ = tse_or_beta_g(g, i + 1, j, k, i, j, k, rho_f, 0);
    }
    if (i != 0) {
        C_ = tse_or_alpha(g, i, j, k, i - 1, j, k, nue, rho_f, Kx, i - 1, i, 0);
        Cnew = C_ * P[IDX(g, i - 1, j, k)];
        D_ = tse_or_beta_g(g, i, j, k, i - 1, j, k, rho_f, 0);
    }
    if (j != ny - 1) {
        E_ = tse_or_alpha(g, i, j + 1, k, i, j, k, nue, rho_f, Ky, j + 1, j, 1);
        Enew = E_ * P[IDX(g, i, j + 1, k)];
    }
    if (j != 0) {
        G_ = tse_or_alpha(g, i, j, k, i, j - 1, k, nue, rho_f, Ky, j - 1, j, 1);
        Gnew = G_ * P[IDX(g, i, j - 1, k)];
    }
    if (k != nz - 1) {
        I_ = tse_or_alpha(g, i, j, k + 1, i, j, k, nue, rho_f, Kz, k + 1, k, 2);
        Inew = I_ * P[IDX(g, i, j, k + 1)];
        J_ = tse_or_beta_g(g, i, j, k + 1, i, j, k, rho_f, 2);
    }
    if (k != 0) {
        K_ = tse_or_alpha(g, i, j, k, i, j, k - 1, nue, rho_f, Kz, k - 1, k, 2);
        Knew = K_ * P[IDX(g, i, j, k - 1)];
        L_ = tse_or_beta_g(g, i, j, k, i, j, k - 1, rho_f, 2);
    }
    const size_t c = IDX(g, i, j, k);
    const double por = phi[c];
    const double R = rho_f[c] * ((1 - por) * c_r[c] + por * c_f[c]) / dt_P;
    const double M = A_ + C_ + E_ + G_ + I_ + K_;
    return (1 - SOR) * P[c] +
           SOR * (Anew + A_ * B_ + Cnew - C_ * D_ + Enew + Gnew + Inew +
                  I_ * J_ + Knew - K_ * L_ + R * Pn[c] + WP[c]) /
               (R + M);
}

void tse_or_p_half_sweep(const tse_or_grid *g, double *P, const double *Pn,
                         const double *c_f, const double *nue,
                         const double *rho_f, const double *phi,
                         const double *Kx, const double *Ky, const double *Kz,
                         const double *BP, const double *WP,
                         const double *c_r, double dt_P, double SOR,
                         int colour) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = ((i + j + colour) & 1); k < nz; k += 2) {
                /* (i+j+k)%2 == colour */
                const size_t c = IDX(g, i, j, k);
                if (!(BP[c] != 0))
                    P[c] = p_solve(g, P, Pn, c_f, nue, rho_f, phi, Kx, Ky, Kz,
                                   WP, c_r, dt_P, SOR, i, j, k);
            }
}

/* Optional progress report of long compute_P calls: called with the
 * iteration count every `g_progress_every` iterations (no effect on any
 * value). */
static void (*g_progress)(int) = 0;
static int g_progress_every = 0;

void tse_or_set_progress(void (*cb)(int), int every) {
    g_progress = cb;
    g_progress_every = every;
}

/* compute_P (tse.py:647-742). */
int tse_or_compute_p(const tse_or_grid *g, double *P, double *Pn,
                     const double *c_f, const double *nue,
                     const double *rho_f, const double *phi,
                     const double *Kx, const double *Ky, const double *Kz,
                     const double *BP, const double *WP, const double *c_r,
                     double dt_P, double SOR, int max_it, double inc_it,
                     double max_dt, double min_dt, double P_conv,
                     tse_or_result *out) {
    const size_t N = (size_t)g->nx * g->ny * g->nz;
    double *Pold = (double *)malloc(N * sizeof(double));
    if (!Pold) return -1;
    double t_P = 0;
    int converged = 0, num_it = 0, num_sub_it = 0, halvings = 0;
    double diff = 1e100;
    memcpy(Pn, P, N * sizeof(double));
    while (!converged) {
        num_it += 1;
        num_sub_it += 1;
        if (g_progress && g_progress_every > 0 && num_it % g_progress_every == 0)
            g_progress(num_it);
        memcpy(Pold, P, N * sizeof(double));
        tse_or_p_half_sweep(g, P, Pn, c_f, nue, rho_f, phi, Kx, Ky, Kz, BP, WP,
                            c_r, dt_P, SOR, 0);
        tse_or_p_half_sweep(g, P, Pn, c_f, nue, rho_f, phi, Kx, Ky, Kz, BP, WP,
                            c_r, dt_P, SOR, 1);
        diff = 1e100;
        if (num_it < 50)
            diff = tse_or_l2_norm(P, Pold, N, NULL);
        else if (num_it % 10 == 0)
            diff = tse_or_l2_norm(P, Pold, N, NULL);
        if (diff < P_conv) {
            converged = 1;
            t_P += dt_P;
            if (num_sub_it < max_it / 2.0) dt_P = py_min(dt_P * inc_it, max_dt);
        } else {
            converged = 0;
            if (num_sub_it >= max_it) {
                dt_P = dt_P / 2;
                num_sub_it = 0;
                halvings += 1;
                memcpy(P, Pn, N * sizeof(double));
            }
            if (dt_P <= min_dt) break;
        }
    }
    free(Pold);
    out->t = t_P;
    out->dt = dt_P;
    out->converged = converged;
    out->num_it = num_it;
    out->num_halvings = halvings;
    out->last_diff = diff;
    return 0;
}

/* ---- compute_q (tse.py:744-780) ----------------------------------------- */
/* smolarkiewicz (tse.py:1083-1116): anti-diffusive velocities of the
 * Smolarkiewicz (1983) correction.  Cells with C > 0 get, per axis with an
 * upper neighbour, Sc*(|q| dd - dt q^2)*(C_nb - C) / ((C + C_nb + eps) dd)
 * with dd the centre distance, evaluated left to right; the last index of
 * an axis keeps its previous value.  Cells with C <= 0 get 0 on all axes.
 * q^2 is q*q (Numba's integer power). */
void tse_or_smolarkiewicz(const tse_or_grid *g, const double *C,
                          const double *qx, const double *qy, const double *qz,
                          double *vx, double *vy, double *vz, double dt,
                          double eps, double Sc) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t c = IDX(g, i, j, k);
                if (C[c] > 0) {
                    if (i < nx - 1) {
                        const size_t e = IDX(g, i + 1, j, k);
                        const double dd = 0.5 * (g->dx[i + 1] + g->dx[i]);
                        const double q = qx[c];
                        vx[c] = Sc * (fabs(q) * dd - dt * (q * q)) * (C[e] - C[c]) /
                                ((C[c] + C[e] + eps) * dd);
                    }
                    if (j < ny - 1) {
                        const size_t e = IDX(g, i, j + 1, k);
                        const double dd = 0.5 * (g->dy[j + 1] + g->dy[j]);
                        const double q = qy[c];
                        vy[c] = Sc * (fabs(q) * dd - dt * (q * q)) * (C[e] - C[c]) /
                                ((C[c] + C[e] + eps) * dd);
                    }
                    if (k < nz - 1) {
                        const size_t e = IDX(g, i, j, k + 1);
                        const double dd = 0.5 * (g->dz[k + 1] + g->dz[k]);
                        const double q = qz[c];
                        vz[c] = Sc * (fabs(q) * dd - dt * (q * q)) * (C[e] - C[c]) /
                                ((C[c] + C[e] + eps) * dd);
                    }
                } else {
                    vx[c] = 0.0;
                    vy[c] = 0.0;
                    vz[c] = 0.0;
                }
            }
}

void tse_or_compute_q(const tse_or_grid *g, const double *P, const double *BP,
                      const double *Kx, const double *Ky, const double *Kz,
                      double *qx, double *qy, double *qz, const double *nue,
                      const double *rho_f) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t c = IDX(g, i, j, k);
                if (i < nx - 1) {
                    const size_t e = IDX(g, i + 1, j, k);
                    if (!(BP[c] != 0 && BP[e] != 0))
                        qx[c] = (-tse_or_harm(g, Kx, i + 1, j, k, i, j, k, 0) /
                                 tse_or_harm(g, nue, i + 1, j, k, i, j, k, 0) *
                                 ((P[e] - P[c]) / (0.5 * (g->dx[i + 1] + g->dx[i])) +
                                  tse_or_amean(g, rho_f, i + 1, j, k, i, j, k, 0) *
                                      g->g * g->sin_beta));
                    else
                        qx[c] = 0;
                }
                if (j < ny - 1) {
                    const size_t n = IDX(g, i, j + 1, k);
                    if (!(BP[c] != 0 && BP[n] != 0))
                        qy[c] = (-tse_or_harm(g, Ky, i, j + 1, k, i, j, k, 1) /
                                 tse_or_harm(g, nue, i, j + 1, k, i, j, k, 1) *
                                 (P[n] - P[c]) / (0.5 * (g->dy[j + 1] + g->dy[j])));
                    else
                        qy[c] = 0;
                }
                if (k < nz - 1) {
                    const size_t u = IDX(g, i, j, k + 1);
                    if (!(BP[c] != 0 && BP[u] != 0))
                        qz[c] = (-tse_or_harm(g, Kz, i, j, k + 1, i, j, k, 2) /
                                 tse_or_harm(g, nue, i, j, k + 1, i, j, k, 2) *
                                 ((P[u] - P[c]) / (0.5 * (g->dz[k + 1] + g->dz[k])) +
                                  tse_or_amean(g, rho_f, i, j, k + 1, i, j, k, 2) *
                                      g->g * g->cos_beta));
                    else
                        qz[c] = 0;
                }
            }
}

/* ---- flow_error (tse.py:409-446) ----------------------------------------- */
void tse_or_flow_error(const tse_or_grid *g, const double *P,
                       const double *Pn, const double *qx, const double *qy,
                       const double *qz, const double *WP, const double *phi,
                       const double *c_f, const double *c_r,
                       const double *rho_f, double dt_P, double *e) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t c = IDX(g, i, j, k);
                double v = 0.0;
                if (i != nx - 1)
                    v -= (qx[c] / g->dx[i] *
                          tse_or_harm(g, rho_f, i + 1, j, k, i, j, k, 0));
                if (i != 0)
                    v += (qx[IDX(g, i - 1, j, k)] / g->dx[i] *
                          tse_or_harm(g, rho_f, i - 1, j, k, i, j, k, 0));
                if (j != ny - 1)
                    v -= (qy[c] / g->dy[j] *
                          tse_or_harm(g, rho_f, i, j + 1, k, i, j, k, 1));
                if (j != 0)
                    v += (qy[IDX(g, i, j - 1, k)] / g->dy[j] *
                          tse_or_harm(g, rho_f, i, j - 1, k, i, j, k, 1));
                if (k != nz - 1)
                    v -= (qz[c] / g->dz[k] *
                          tse_or_harm(g, rho_f, i, j, k + 1, i, j, k, 2));
                if (k != 0)
                    v += (qz[IDX(g, i, j, k - 1)] / g->dz[k] *
                          tse_or_harm(g, rho_f, i, j, k - 1, i, j, k, 2));
                const double por = phi[c];
                v += (WP[c] + ((1 - por) * c_r[c] + por * c_f[c]) *
                                  (P[c] - Pn[c]) * rho_f[c] / dt_P);
                e[c] = v;
            }
}

/* ---- compute_C ------------------------------------------------------------ */

/* solve() of compute_C (tse.py:795-875). */
static inline double c_solve(const tse_or_grid *g, const double *C,
                             const double *Cn, const double *D_fx,
                             const double *D_fy, const double *D_fz,
                             const double *phi, const double *rho_f,
                             const double *qx, const double *qy,
                             const double *qz, const double *BC,
                             const double *WC, const double *RC,
                             const double *e, const double *lf_l,
                             const double *lf_t, int imblflo, double dt_C,
                             double omega, double SOR, int i, int j, int k) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double A_ = 0, B_ = 0, C_ = 0, D_ = 0, E_ = 0, F_ = 0, G_ = 0, H_ = 0,
           I_ = 0, J_ = 0, K_ = 0, L_ = 0, M_ = 0, N_ = 0, O_ = 0, P_ = 0,
           Q_ = 0, R_ = 0, EC = 0, Mnew = 0, Nnew = 0, Onew = 0, Pnew = 0,
           Qnew = 0, Rnew = 0, Mold = 0, Nold = 0, Oold = 0, Pold = 0,
           Qold = 0, Rold = 0;
    double D[6] = {0, 0, 0, 0, 0, 0}; /* D_xl, D_yl, D_zl, D_xr, D_yr, D_zr */
    const size_t c = IDX(g, i, j, k);
    if (lf_l[c] != 0 || lf_t[c] != 0)
        tse_or_dispersivity(g, qx, qy, qz, lf_l, lf_t, BC, i, j, k, D);
    double dd, dd2, ph;
    if (i != nx - 1) {
        const size_t o = IDX(g, i + 1, j, k);
        dd = g->dx[i];
        dd2 = (0.5 * (g->dx[i + 1] + g->dx[i])) * dd;
        ph = tse_or_amean(g, phi, i + 1, j, k, i, j, k, 0);
        A_ = (tse_or_harm(g, D_fx, i + 1, j, k, i, j, k, 0) * ph + D[3]) / dd2;
        N_ = py_min0(qx[c]) / dd;
        H_ = py_max0(qx[c]) / dd;
        Mnew = (A_ - N_) * C[o];
        Mold = (A_ - N_) * Cn[o];
    }
    if (i != 0) {
        const size_t o = IDX(g, i - 1, j, k);
        dd = g->dx[i];
        dd2 = (0.5 * (g->dx[i - 1] + g->dx[i])) * dd;
        ph = tse_or_amean(g, phi, i - 1, j, k, i, j, k, 0);
        B_ = (tse_or_harm(g, D_fx, i - 1, j, k, i, j, k, 0) * ph + D[0]) / dd2;
        M_ = py_max0(qx[o]) / dd;
        G_ = py_min0(qx[o]) / dd;
        Nnew = (B_ + M_) * C[o];
        Nold = (B_ + M_) * Cn[o];
    }
    if (j != ny - 1) {
        const size_t o = IDX(g, i, j + 1, k);
        dd = g->dy[j];
        dd2 = (0.5 * (g->dy[j + 1] + g->dy[j])) * dd;
        ph = tse_or_amean(g, phi, i, j + 1, k, i, j, k, 1);
        C_ = (tse_or_harm(g, D_fy, i, j + 1, k, i, j, k, 1) * ph + D[4]) / dd2;
        P_ = py_min0(qy[c]) / dd;
        J_ = py_max0(qy[c]) / dd;
        Onew = (C_ - P_) * C[o];
        Oold = (C_ - P_) * Cn[o];
    }
    if (j != 0) {
        const size_t o = IDX(g, i, j - 1, k);
        dd = g->dy[j];
        dd2 = (0.5 * (g->dy[j - 1] + g->dy[j])) * dd;
        ph = tse_or_amean(g, phi, i, j - 1, k, i, j, k, 1);
        D_ = (tse_or_harm(g, D_fy, i, j - 1, k, i, j, k, 1) * ph + D[1]) / dd2;
        O_ = py_max0(qy[o]) / dd;
        I_ = py_min0(qy[o]) / dd;
        Pnew = (D_ + O_) * C[o];
        Pold = (D_ + O_) * Cn[o];
    }
    if (k != nz - 1) {
        const size_t o = IDX(g, i, j, k + 1);
        dd = g->dz[k];
        dd2 = (0.5 * (g->dz[k + 1] + g->dz[k])) * dd;
        ph = tse_or_amean(g, phi, i, j, k + 1, i, j, k, 2);
        E_ = (tse_or_harm(g, D_fz, i, j, k + 1, i, j, k, 2) * ph + D[5]) / dd2;
        R_ = py_min0(qz[c]) / dd;
        L_ = py_max0(qz[c]) / dd;
        Qnew = (E_ - R_) * C[o];
        Qold = (E_ - R_) * Cn[o];
    }
    if (k != 0) {
        const size_t o = IDX(g, i, j, k - 1);
        dd = g->dz[k];
        dd2 = (0.5 * (g->dz[k - 1] + g->dz[k])) * dd;
        ph = tse_or_amean(g, phi, i, j, k - 1, i, j, k, 2);
        F_ = (tse_or_harm(g, D_fz, i, j, k - 1, i, j, k, 2) * ph + D[2]) / dd2;
        Q_ = py_max0(qz[o]) / dd;
        K_ = py_min0(qz[o]) / dd;
        Rnew = (F_ + Q_) * C[o];
        Rold = (F_ + Q_) * Cn[o];
    }
    const double R = RC[c] * phi[c] / dt_C;
    const double M = -A_ - B_ - C_ - D_ - E_ - F_ + G_ - H_ + I_ - J_ + K_ - L_;
    if (imblflo) EC = Cn[c] / rho_f[c] * e[c];
    return ((1 - SOR) * C[c] +
            SOR *
                (omega * (Mnew + Nnew + Onew + Pnew + Qnew + Rnew) +
                 (1 - omega) *
                     (Cn[c] * M + Mold + Nold + Oold + Pold + Qold + Rold) +
                 R * Cn[c] + WC[c] - EC) /
                (R - omega * M));
}

int tse_or_compute_c(const tse_or_grid *g, double *C, double *Cn_out,
                     const double *D_fx, const double *D_fy,
                     const double *D_fz, const double *phi,
                     const double *rho_f, const double *qx, const double *qy,
                     const double *qz, const double *BC, const double *WC,
                     const double *RC, const double *e, const double *lf_l,
                     const double *lf_t, int imblflo, double dt_C,
                     double omega, double SOR, int max_it, double inc_it,
                     double max_dt, double min_dt, double C_conv,
                     tse_or_result *out) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t N = (size_t)nx * ny * nz;
    double *Cn = Cn_out;
    double *Cold = (double *)malloc(N * sizeof(double));
    if (!Cold) return -1;
    double t_C = 0, diff = 0;
    int converged = 0, num_it = 0, num_sub_it = 0, halvings = 0;
    memcpy(Cn, C, N * sizeof(double));
    while (!converged) {
        num_it += 1;
        num_sub_it += 1;
        memcpy(Cold, C, N * sizeof(double));
        for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int k = ((i + j + colour) & 1); k < nz; k += 2) {
                        const size_t c = IDX(g, i, j, k);
                        if (!(BC[c] != 0))
                            C[c] = c_solve(g, C, Cn, D_fx, D_fy, D_fz, phi,
                                           rho_f, qx, qy, qz, BC, WC, RC, e,
                                           lf_l, lf_t, imblflo, dt_C, omega,
                                           SOR, i, j, k);
                    }
        }
        diff = tse_or_l2_norm(C, Cold, N, NULL);
        if (diff < C_conv || omega == 0) {
            t_C += dt_C;
            if (num_sub_it < max_it / 2.0) dt_C = py_min(dt_C * inc_it, max_dt);
            converged = 1;
        } else {
            converged = 0;
            if (num_sub_it >= max_it) {
                dt_C = dt_C / 2;
                num_sub_it = 0;
                halvings += 1;
                memcpy(C, Cn, N * sizeof(double));
            }
            if (dt_C <= min_dt) break;
        }
    }
    free(Cold);
    out->t = t_C;
    out->dt = dt_C;
    out->converged = converged;
    out->num_it = num_it;
    out->num_halvings = halvings;
    out->last_diff = diff;
    return 0;
}

/* ---- compute_T ------------------------------------------------------------ */

/* One face of compute_T's solve() (tse.py:943-1031): returns the face
 * conductance term and S_ = harm(rho_f)*harm(cp_f)/dd. */
static inline void t_face(const tse_or_grid *g, const double *phi,
                          const double *l_f, const double *l_r,
                          const double *cp_f, const double *rho_f, int i1,
                          int j1, int k1, int i2, int j2, int k2, int n,
                          double dd, double dd2, double Ddisp, double *A,
                          double *S) {
    const double por = tse_or_amean(g, phi, i1, j1, k1, i2, j2, k2, n);
    const double lf = tse_or_harm(g, l_f, i1, j1, k1, i2, j2, k2, n);
    const double lr = tse_or_harm(g, l_r, i1, j1, k1, i2, j2, k2, n);
    const double cpf = tse_or_harm(g, cp_f, i1, j1, k1, i2, j2, k2, n);
    const double rhocpf = tse_or_harm(g, rho_f, i1, j1, k1, i2, j2, k2, n) * cpf;
    *A = ((1 - por) * lr + por * lf + Ddisp * rhocpf) / dd2;
    *S = rhocpf / dd;
}

/* solve() of compute_T (tse.py:930-1043). */
static inline double t_solve(const tse_or_grid *g, const double *T,
                             const double *Tn, const double *l_f,
                             const double *cp_f, const double *rho_f,
                             const double *phi, const double *qx,
                             const double *qy, const double *qz,
                             const double *BT, const double *WT,
                             const double *e, const double *l_r,
                             const double *cp_r, const double *rho_r,
                             const double *lt_l, const double *lt_t,
                             int imblflo, double dt_T, double omega,
                             double SOR, int i, int j, int k) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double A_ = 0, B_ = 0, C_ = 0, D_ = 0, E_ = 0, F_ = 0, G_ = 0, H_ = 0,
           I_ = 0, J_ = 0, K_ = 0, L_ = 0, S_ = 0, M_ = 0, N_ = 0, O_ = 0,
           P_ = 0, Q_ = 0, R_ = 0, ET = 0, Mnew = 0, Nnew = 0, Onew = 0,
           Pnew = 0, Qnew = 0, Rnew = 0, Mold = 0, Nold = 0, Oold = 0,
           Pold = 0, Qold = 0, Rold = 0;
    double D[6] = {0, 0, 0, 0, 0, 0};
    const size_t c = IDX(g, i, j, k);
    if (lt_l[c] != 0 || lt_t[c] != 0)
        tse_or_dispersivity(g, qx, qy, qz, lt_l, lt_t, BT, i, j, k, D);
    if (i != nx - 1) {
        const size_t o = IDX(g, i + 1, j, k);
        const double dd = g->dx[i];
        const double dd2 = (0.5 * (g->dx[i + 1] + g->dx[i])) * dd;
        t_face(g, phi, l_f, l_r, cp_f, rho_f, i, j, k, i + 1, j, k, 0, dd, dd2,
               D[3], &A_, &S_);
        N_ = py_min0(qx[c]) * S_;
        H_ = py_max0(qx[c]) * S_;
        Mnew = (A_ - N_) * T[o];
        Mold = (A_ - N_) * Tn[o];
    }
    if (i != 0) {
        const size_t o = IDX(g, i - 1, j, k);
        const double dd = g->dx[i];
        const double dd2 = (0.5 * (g->dx[i - 1] + g->dx[i])) * dd;
        t_face(g, phi, l_f, l_r, cp_f, rho_f, i, j, k, i - 1, j, k, 0, dd, dd2,
               D[0], &B_, &S_);
        M_ = py_max0(qx[o]) * S_;
        G_ = py_min0(qx[o]) * S_;
        Nnew = (B_ + M_) * T[o];
        Nold = (B_ + M_) * Tn[o];
    }
    if (j != ny - 1) {
        const size_t o = IDX(g, i, j + 1, k);
        const double dd = g->dy[j];
        const double dd2 = (0.5 * (g->dy[j + 1] + g->dy[j])) * dd;
        t_face(g, phi, l_f, l_r, cp_f, rho_f, i, j + 1, k, i, j, k, 1, dd, dd2,
               D[4], &C_, &S_);
        P_ = py_min0(qy[c]) * S_;
        J_ = py_max0(qy[c]) * S_;
        Onew = (C_ - P_) * T[o];
        Oold = (C_ - P_) * Tn[o];
    }
    if (j != 0) {
        const size_t o = IDX(g, i, j - 1, k);
        const double dd = g->dy[j];
        const double dd2 = (0.5 * (g->dy[j - 1] + g->dy[j])) * dd;
        t_face(g, phi, l_f, l_r, cp_f, rho_f, i, j - 1, k, i, j, k, 1, dd, dd2,
               D[1], &D_, &S_);
        O_ = py_max0(qy[o]) * S_;
        I_ = py_min0(qy[o]) * S_;
        Pnew = (D_ + O_) * T[o];
        Pold = (D_ + O_) * Tn[o];
    }
    if (k != nz - 1) {
        const size_t o = IDX(g, i, j, k + 1);
        const double dd = g->dz[k];
        const double dd2 = (0.5 * (g->dz[k + 1] + g->dz[k])) * dd;
        t_face(g, phi, l_f, l_r, cp_f, rho_f, i, j, k + 1, i, j, k, 2, dd, dd2,
               D[5], &E_, &S_);
        R_ = py_min0(qz[c]) * S_;
        L_ = py_max0(qz[c]) * S_;
        Qnew = (E_ - R_) * T[o];
        Qold = (E_ - R_) * Tn[o];
    }
    if (k != 0) {
        const size_t o = IDX(g, i, j, k - 1);
        const double dd = g->dz[k];
        const double dd2 = (0.5 * (g->dz[k - 1] + g->dz[k])) * dd;
        t_face(g, phi, l_f, l_r, cp_f, rho_f, i, j, k - 1, i, j, k, 2, dd, dd2,
               D[2], &F_, &S_);
        Q_ = py_max0(qz[o]) * S_;
        K_ = py_min0(qz[o]) * S_;
        Rnew = (F_ + Q_) * T[o];
        Rold = (F_ + Q_) * Tn[o];
    }
    const double por_ = phi[c];
    const double R =
        (por_ * rho_f[c] * cp_f[c] + (1 - por_) * rho_r[c] * cp_r[c]) / dt_T;
    const double M = -A_ - B_ - C_ - D_ - E_ - F_ + G_ - H_ + I_ - J_ + K_ - L_;
    if (imblflo) ET = cp_f[c] * Tn[c] * e[c];
    return ((1 - SOR) * T[c] +
            SOR *
                (omega * (Mnew + Nnew + Onew + Pnew + Qnew + Rnew) +
                 (1 - omega) *
                     (Tn[c] * M + Mold + Nold + Oold + Pold + Qold + Rold) +
                 R * Tn[c] + WT[c] - ET) /
                (R - omega * M));
}

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
                     tse_or_result *out) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t N = (size_t)nx * ny * nz;
    double *Tn = Tn_out;
    double *Told = (double *)malloc(N * sizeof(double));
    if (!Told) return -1;
    double t_T = 0, diff = 0;
    int converged = 0, num_it = 0, num_sub_it = 0, halvings = 0;
    memcpy(Tn, T, N * sizeof(double));
    while (!converged) {
        num_it += 1;
        num_sub_it += 1;
        memcpy(Told, T, N * sizeof(double));
        for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int k = ((i + j + colour) & 1); k < nz; k += 2) {
                        const size_t c = IDX(g, i, j, k);
                        if (!(BT[c] != 0))
                            T[c] = t_solve(g, T, Tn, l_f, cp_f, rho_f, phi, qx,
                                           qy, qz, BT, WT, e, l_r, cp_r, rho_r,
                                           lt_l, lt_t, imblflo, dt_T, omega,
                                           SOR, i, j, k);
                    }
        }
        diff = tse_or_l2_norm(T, Told, N, NULL);
        if (diff < T_conv || omega == 0) {
            t_T += dt_T;
            if (num_sub_it < max_it / 2.0) dt_T = py_min(dt_T * inc_it, max_dt);
            converged = 1;
        } else {
            converged = 0;
            if (num_sub_it >= max_it) {
                dt_T = dt_T / 2;
                num_sub_it = 0;
                halvings += 1;
                memcpy(T, Tn, N * sizeof(double));
            }
            if (dt_T <= min_dt) break;
        }
    }
    free(Told);
    out->t = t_T;
    out->dt = dt_T;
    out->converged = converged;
    out->num_it = num_it;
    out->num_halvings = halvings;
    out->last_diff = diff;
    return 0;
}
