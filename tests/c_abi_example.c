/* A plain-C host of the C-ABI (include/tse_hip.h): what a non-Python caller
 * of compute_q / compute_P does.  Device memory comes from the HIP runtime;
 * fields go through tse_to_split / tse_from_split in the reference layout
 * (C order (nx, ny, nz)).
 *
 * 1. compute_q (tse.py:744-780) on P = P0 + a x - rho g z with beta = 0,
 *    uniform K, nue, rho: Darcy's law gives qx = -K/nue * a exactly and
 *    qz = 0 (the z gradient balances gravity).
 * 2. compute_P (tse.py:647-742), incompressible, from a perturbed start
 *    with Dirichlet hydrostatic x faces: the solve must converge back to
 *    the hydrostatic column, the fixed point of the discrete operator.
 * Prints one line per check and exits 0 only if both pass. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "tse_hip.h"

#define CHECK(x)                                                             \
    do {                                                                     \
        int rc_ = (x);                                                       \
        if (rc_) {                                                           \
            fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, tse_last_error()); \
            return 1;                                                        \
        }                                                                    \
    } while (0)
#define HCHECK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                        \
        }                                                                    \
    } while (0)

enum { NX = 12, NY = 7, NZ = 9, N = NX * NY * NZ, NF = 14 };
static const double G = 9.80665, D = 10.0, K = 2e-12, NUE = 1e-3, RHO = 1000.0;
static const double P0 = 2e5, A = -30.0; /* Pa per m along x */

static size_t at(int i, int j, int k) { return ((size_t)i * NY + j) * NZ + k; }

/* host (C order) -> split-layout device field, through a device staging buffer */
static int upload(tse_ctx *ctx, double *dst, const double *host, double *stage) {
    HCHECK(hipMemcpy(stage, host, N * sizeof(double), hipMemcpyHostToDevice));
    CHECK(tse_to_split(ctx, dst, stage, NY * NZ, NZ, 1));
    return 0;
}

static int download(tse_ctx *ctx, double *host, const double *src, double *stage) {
    CHECK(tse_from_split(ctx, stage, NY * NZ, NZ, 1, src));
    HCHECK(hipMemcpy(host, stage, N * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int main(void) {
    double dx[NX], dy[NY], dz[NZ];
    for (int i = 0; i < NX; ++i) dx[i] = D;
    for (int j = 0; j < NY; ++j) dy[j] = D;
    for (int k = 0; k < NZ; ++k) dz[k] = D;
    if (tse_abi_version() != TSE_ABI_VERSION) {
        fprintf(stderr, "ABI version mismatch\n");
        return 1;
    }
    tse_ctx *ctx;
    CHECK(tse_ctx_create(&ctx, 0, NULL));
    CHECK(tse_set_grid(ctx, NX, NY, NZ, dx, dy, dz, G, 0.0, 1.0)); /* beta = 0 */
    const size_t F = (size_t)tse_field_elems(ctx);

    /* device fields: P Pn c_f nue rho phi K BP WP c_r qx qy qz e */
    double *d[NF], *stage;
    for (int f = 0; f < NF; ++f) HCHECK(hipMalloc((void **)&d[f], F * sizeof(double)));
    HCHECK(hipMalloc((void **)&stage, N * sizeof(double)));
    double *h = malloc(N * sizeof(double)), *h2 = malloc(N * sizeof(double));
    if (!h || !h2) return 1;

    /* incompressible (c_f = c_r = 0): no storage term, so the hydrostatic
       column is the fixed point whatever the start */
    const double uni[][2] = {{2, 0.0}, {3, NUE}, {4, RHO}, {5, 0.2}, {6, K},
                             {8, 0.0}, {9, 0.0}};
    for (size_t u = 0; u < sizeof(uni) / sizeof(uni[0]); ++u) {
        for (int c = 0; c < N; ++c) h[c] = uni[u][1];
        if (upload(ctx, d[(int)uni[u][0]], h, stage)) return 1;
    }
    /* 1. compute_q on a linear pressure with hydrostatic z gradient */
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
                h[at(i, j, k)] = P0 + A * (i + 0.5) * D - RHO * G * (k + 0.5) * D;
    if (upload(ctx, d[0], h, stage)) return 1;
    for (int c = 0; c < N; ++c) h2[c] = 0.0;                       /* BP = 0 */
    if (upload(ctx, d[7], h2, stage)) return 1;
    for (int f = 10; f < 13; ++f)
        if (upload(ctx, d[f], h2, stage)) return 1;
    CHECK(tse_compute_q(ctx, d[0], d[7], d[6], d[6], d[6], d[10], d[11], d[12],
                        d[3], d[4]));
    double qx_err = 0.0, qz_max = 0.0;
    const double qx_exact = -K / NUE * A;
    if (download(ctx, h, d[10], stage)) return 1;
    for (int i = 0; i < NX - 1; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
                qx_err = fmax(qx_err, fabs(h[at(i, j, k)] - qx_exact));
    if (download(ctx, h, d[12], stage)) return 1;
    for (int c = 0; c < N; ++c) qz_max = fmax(qz_max, fabs(h[c]));
    /* qz: the z gradient cancels gravity; rounding leaves a tiny fraction of
       the gravity-driven flux scale K/nue*rho*g */
    const int q_ok = qx_err <= 1e-12 * fabs(qx_exact) &&
                     qz_max <= 1e-12 * K / NUE * RHO * G;
    printf("compute_q: qx error %.3e of %.3e, max |qz| %.3e: %s\n", qx_err,
           fabs(qx_exact), qz_max, q_ok ? "ok" : "FAIL");

    /* 2. compute_P back to the hydrostatic column */
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k) {
                const size_t c = at(i, j, k);
                const int dir = (i == 0 || i == NX - 1);
                h[c] = P0 - RHO * G * (k + 0.5) * D +
                       (dir ? 0.0 : 50.0 * sin(0.7 * i + 1.3 * j + 0.9 * k));
                h2[c] = dir ? 1.0 : 0.0;
            }
    if (upload(ctx, d[0], h, stage)) return 1;
    if (upload(ctx, d[7], h2, stage)) return 1;
    tse_p_fields pf = {d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[6], d[6], d[7],
                       d[8], d[9]};
    tse_sor_params prm = {86400.0, 1.8, 1.0, 1.25, 86400.0, 1e-8, 1e-13, 200000,
                          TSE_FORM_FAITHFUL64};
    tse_result res;
    CHECK(tse_compute_p(ctx, &pf, &prm, &res));
    if (download(ctx, h, d[0], stage)) return 1;
    double p_err = 0.0;
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k)
                p_err = fmax(p_err, fabs(h[at(i, j, k)] - (P0 - RHO * G * (k + 0.5) * D)));
    const int p_ok = res.converged && p_err <= 1e-6;
    printf("compute_P: converged %d after %d iterations, max |P - hydrostatic| "
           "%.3e Pa: %s\n", res.converged, res.num_it, p_err, p_ok ? "ok" : "FAIL");

    for (int f = 0; f < NF; ++f) (void)hipFree(d[f]);
    (void)hipFree(stage);
    free(h);
    free(h2);
    CHECK(tse_ctx_destroy(ctx));
    return q_ok && p_ok ? 0 : 1;
}
