// tse_step.hip -- device-resident pieces of the SIA step around the solvers:
// the closed-form linear EOS / conductivity update of the Picard loop
// (tse.py:1422-1423 calls config.fluid_properties) and FP64 mass-balance
// sums (tse.py:1461-1463 calls config.mass_balance).  Both act on
// split-layout fields of one context.
#include <algorithm>
#include <cmath>

#include "tse_internal.h"

namespace tse {

__global__ void __launch_bounds__(kBlock)
linear_eos_kernel(Geo G, const double *__restrict__ C, const double *__restrict__ T,
                  double *__restrict__ rho, double *__restrict__ l_f,
                  tse_eos_params p) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        rho[t] = p.rho0 + p.drho_dc * C[t];
        if (l_f) l_f[t] = p.lf0 * (1.0 + p.dlf_dt * (T[t] - p.t_ref));
    }
}

__device__ inline double wsum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Per-block partial sums: [sum C*phi*V, sum rho*phi*V, outflow rate], the
// outflow being the upwind flux of C through the two x faces: at i = nx-2
// (flux qx stored at the lower cell of the last face) max(qx,0)*C, and at
// i = 0 max(-qx,0)*C of cell 1, each times the face area dy*dz.
__global__ void __launch_bounds__(kBlock)
mass_kernel(Geo G, const double *__restrict__ dxyz, const double *__restrict__ C,
            const double *__restrict__ rho, const double *__restrict__ phi,
            const double *__restrict__ qx, double *__restrict__ partials) {
    const double *dx = dxyz, *dy = dxyz + G.nx, *dz = dxyz + G.nx + G.ny;
    const long F = 2L * G.S;
    double a = 0, b = 0, o = 0;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        if (!q.valid) continue;
        const double V = dx[q.i] * dy[q.j] * dz[q.k];
        a += C[t] * phi[t] * V;
        b += rho[t] * phi[t] * V;
        if (G.nx >= 3) {
            const double area = dy[q.j] * dz[q.k];
            if (q.i == G.nx - 2) o += fmax(qx[t], 0.0) * C[t] * area;
            if (q.i == 0) {
                const long t1 = slot(G, 1, q.j, q.k);
                o += fmax(-qx[t], 0.0) * C[t1] * area;
            }
        }
    }
    __shared__ double s[3][kBlock / 64];
    a = wsum_d(a);
    b = wsum_d(b);
    o = wsum_d(o);
    if ((threadIdx.x & 63) == 0) {
        s[0][threadIdx.x >> 6] = a;
        s[1][threadIdx.x >> 6] = b;
        s[2][threadIdx.x >> 6] = o;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0, z = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            x += s[0][w];
            y += s[1][w];
            z += s[2][w];
        }
        partials[3 * blockIdx.x] = x;
        partials[3 * blockIdx.x + 1] = y;
        partials[3 * blockIdx.x + 2] = z;
    }
}

__global__ void __launch_bounds__(kBlock)
mass_reduce_kernel(const double *__restrict__ partials, int n, double *__restrict__ out) {
    double v[3] = {0, 0, 0};
    for (int q = threadIdx.x; q < n; q += blockDim.x)
        for (int c = 0; c < 3; ++c) v[c] += partials[3 * q + c];
    __shared__ double s[3][kBlock / 64];
    for (int c = 0; c < 3; ++c) {
        const double w = wsum_d(v[c]);
        if ((threadIdx.x & 63) == 0) s[c][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) {
            double x = 0;
            for (int w = 0; w < kBlock / 64; ++w) x += s[c][w];
            out[c] = x;
        }
}

// ---- Peclet numbers and CFL maxima (diagnostics of the SIA driver) ----------------------
// numpy.max (and Numba's array max) return NaN when any element is NaN.
__device__ inline double max_nan(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}

__device__ inline double wmax_nan(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max_nan(v, __shfl_xor(v, o, 64));
    return v;
}

// Block maxima of (m0, m1, m2) to partials[3*block + a].
__device__ inline void block_max3(double m0, double m1, double m2, double *partials) {
    __shared__ double s[3][kBlock / 64];
    m0 = wmax_nan(m0);
    m1 = wmax_nan(m1);
    m2 = wmax_nan(m2);
    if ((threadIdx.x & 63) == 0) {
        s[0][threadIdx.x >> 6] = m0;
        s[1][threadIdx.x >> 6] = m1;
        s[2][threadIdx.x >> 6] = m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v[3] = {s[0][0], s[1][0], s[2][0]};
        for (int w = 1; w < kBlock / 64; ++w)
            for (int a = 0; a < 3; ++a) v[a] = max_nan(v[a], s[a][w]);
        for (int a = 0; a < 3; ++a) partials[3 * blockIdx.x + a] = v[a];
    }
}

__global__ void __launch_bounds__(kBlock)
max3_reduce_kernel(const double *__restrict__ partials, int n, double *__restrict__ out) {
    double v[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int q = threadIdx.x; q < n; q += blockDim.x)
        for (int a = 0; a < 3; ++a) v[a] = max_nan(v[a], partials[3 * q + a]);
    block_max3(v[0], v[1], v[2], out);   // one block: out[0..2]
}

// C-order cell f of (nx, ny, nz) -> (i, j, k).
__device__ inline void unflat(const Geo &G, long f, int &i, int &j, int &k) {
    k = (int)(f % G.nz);
    const long r = f / G.nz;
    j = (int)(r % G.ny);
    i = (int)(r / G.ny);
}

// peclet_C (tse.py:333-360) / peclet_T (tse.py:364-383).  Pe and the
// species-last D arrays are in the host layout; q, B and the heat fields are
// split-layout fields.  Cells with B != 0 keep their Pe.
template <int HEAT>
__global__ void __launch_bounds__(kBlock)
peclet_kernel(Geo G, const double *__restrict__ dxyz, const double *__restrict__ qx,
              const double *__restrict__ qy, const double *__restrict__ qz,
              const double *__restrict__ Dx, const double *__restrict__ Dy,
              const double *__restrict__ Dz, int nspec, const double *__restrict__ rho,
              const double *__restrict__ cp, const double *__restrict__ phi,
              const double *__restrict__ l_r, const double *__restrict__ l_f,
              const double *__restrict__ B, double *__restrict__ Pe,
              double *__restrict__ partials) {
    const double *dx = dxyz, *dy = dxyz + G.nx, *dz = dxyz + G.nx + G.ny;
    double m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY;
    // peclet_T reads nine split-layout fields per cell and writes three host
    // doubles: cells go in split-slot order (x fastest).  peclet_C reads
    // 3*nspec host-layout D values per cell: host order (k fastest).  The
    // maxima do not depend on the order.
    const long nit = HEAT ? 2L * G.S : (long)G.nx * G.ny * G.nz;
    for (long u = blockIdx.x * (long)blockDim.x + threadIdx.x; u < nit;
         u += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        long f, t;
        if (HEAT) {
            const int c = u >= G.S ? 1 : 0;
            const Cell q = decode(G, c, (int)(u - (long)c * G.S));
            if (!q.valid) continue;
            i = q.i;
            j = q.j;
            k = q.k;
            f = ((long)i * G.ny + j) * G.nz + k;
            t = u;
        } else {
            f = u;
            unflat(G, f, i, j, k);
            t = slot(G, i, j, k);
        }
        double v0, v1, v2;
        if (B[t] == 0) {
            if (HEAT) {
                const double den = (1 - phi[t]) * l_r[t] + phi[t] * l_f[t];
                v0 = fabs(qx[t]) * rho[t] * cp[t] / den * dx[i];
                v1 = fabs(qy[t]) * rho[t] * cp[t] / den * dy[j];
                v2 = fabs(qz[t]) * rho[t] * cp[t] / den * dz[k];
            } else {
                double mx = Dx[f * nspec], my = Dy[f * nspec], mz = Dz[f * nspec];
                for (int q = 1; q < nspec; ++q) {
                    mx = max_nan(mx, Dx[f * nspec + q]);
                    my = max_nan(my, Dy[f * nspec + q]);
                    mz = max_nan(mz, Dz[f * nspec + q]);
                }
                v0 = mx != 0 ? fabs(qx[t]) / mx * dx[i] : NAN;
                v1 = my != 0 ? fabs(qy[t]) / my * dy[j] : NAN;
                v2 = mz != 0 ? fabs(qz[t]) / mz * dz[k] : NAN;
            }
            Pe[3 * f] = v0;
            Pe[3 * f + 1] = v1;
            Pe[3 * f + 2] = v2;
        } else {
            v0 = Pe[3 * f];
            v1 = Pe[3 * f + 1];
            v2 = Pe[3 * f + 2];
        }
        m0 = max_nan(m0, v0);
        m1 = max_nan(m1, v1);
        m2 = max_nan(m2, v2);
    }
    block_max3(m0, m1, m2, partials);
}

// CFL maxima of the SIA driver (tse.py:1241-1246): |qx|/phi is reshaped to
// (ny, nz, nx) and |qy|/phi to (nx, nz, ny) before the division by dx / dy,
// so C-order element f is divided by dx[f % nx], dy[f % ny] and dz[f % nz].
__global__ void __launch_bounds__(kBlock)
cfl_kernel(Geo G, const double *__restrict__ dxyz, const double *__restrict__ qx,
           const double *__restrict__ qy, const double *__restrict__ qz,
           const double *__restrict__ phi, double *__restrict__ partials) {
    const double *dx = dxyz, *dy = dxyz + G.nx, *dz = dxyz + G.nx + G.ny;
    double m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY;
    // every operand is a split-layout field: cells in split-slot order, each
    // with its host C-order index f for the reshaped spacing
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 2L * G.S;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        if (!q.valid) continue;
        const int k = q.k;
        const long f = ((long)q.i * G.ny + q.j) * G.nz + k;
        m0 = max_nan(m0, fabs(qx[t]) / phi[t] / dx[f % G.nx]);
        m1 = max_nan(m1, fabs(qy[t]) / phi[t] / dy[f % G.ny]);
        m2 = max_nan(m2, fabs(qz[t]) / phi[t] / dz[k]);
    }
    block_max3(m0, m1, m2, partials);
}

static int finish_max3(tse_ctx *ctx, int nb, double *out3) {
    hipLaunchKernelGGL(max3_reduce_kernel, dim3(1), dim3(kBlock), 0, ctx->stream,
                       ctx->partials, nb, ctx->red_out);
    TSE_HIP_TRY(hipGetLastError());
    TSE_HIP_TRY(hipMemcpyAsync(ctx->red_host, ctx->red_out, 3 * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int a = 0; a < 3; ++a) out3[a] = ctx->red_host[a];
    return 0;
}

static int diag_blocks(const Geo &G) {
    const long N = (long)G.nx * G.ny * G.nz;
    return (int)std::max<long>(1, std::min<long>((N + kBlock - 1) / kBlock, 1024));
}

}  // namespace tse

using namespace tse;

extern "C" {

int tse_peclet_c(tse_ctx *ctx, const double *qx, const double *qy, const double *qz,
                 const double *D_fx, const double *D_fy, const double *D_fz,
                 int32_t nspec, const double *BC, double *PeC, double *out3) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!qx || !qy || !qz || !D_fx || !D_fy || !D_fz || !BC || !PeC || !out3)
        return fail(TSE_ERR_ARG, "null argument");
    if (nspec < 1) return fail(TSE_ERR_ARG, "nspec must be >= 1");
    const Geo &G = ctx->geo;
    const int nb = diag_blocks(G);
    hipLaunchKernelGGL((peclet_kernel<0>), dim3(nb), dim3(kBlock), 0, ctx->stream, G,
                       ctx->d_dxyz, qx, qy, qz, D_fx, D_fy, D_fz, (int)nspec, nullptr,
                       nullptr, nullptr, nullptr, nullptr, BC, PeC, ctx->partials);
    TSE_HIP_TRY(hipGetLastError());
    return finish_max3(ctx, nb, out3);
}

int tse_peclet_t(tse_ctx *ctx, const double *qx, const double *qy, const double *qz,
                 const double *rho_f, const double *cp_f, const double *phi,
                 const double *l_r, const double *l_f, const double *BT, double *PeT,
                 double *out3) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!qx || !qy || !qz || !rho_f || !cp_f || !phi || !l_r || !l_f || !BT || !PeT ||
        !out3)
        return fail(TSE_ERR_ARG, "null argument");
    const Geo &G = ctx->geo;
    const int nb = diag_blocks(G);
    hipLaunchKernelGGL((peclet_kernel<1>), dim3(nb), dim3(kBlock), 0, ctx->stream, G,
                       ctx->d_dxyz, qx, qy, qz, nullptr, nullptr, nullptr, 1, rho_f,
                       cp_f, phi, l_r, l_f, BT, PeT, ctx->partials);
    TSE_HIP_TRY(hipGetLastError());
    return finish_max3(ctx, nb, out3);
}

int tse_cfl_maxima(tse_ctx *ctx, const double *qx, const double *qy, const double *qz,
                   const double *phi, double *out3) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!qx || !qy || !qz || !phi || !out3) return fail(TSE_ERR_ARG, "null argument");
    const Geo &G = ctx->geo;
    const int nb = diag_blocks(G);
    hipLaunchKernelGGL(cfl_kernel, dim3(nb), dim3(kBlock), 0, ctx->stream, G,
                       ctx->d_dxyz, qx, qy, qz, phi, ctx->partials);
    TSE_HIP_TRY(hipGetLastError());
    return finish_max3(ctx, nb, out3);
}

int tse_linear_eos(tse_ctx *ctx, const double *C, const double *T, double *rho_f,
                   double *l_f, const tse_eos_params *p) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!C || !rho_f || !p || (l_f && !T)) return fail(TSE_ERR_ARG, "null argument");
    const Geo &G = ctx->geo;
    const long F = 2L * G.S;
    const int nb = (int)std::min<long>((F + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(linear_eos_kernel, dim3(nb), dim3(kBlock), 0, ctx->stream, G,
                       C, T, rho_f, l_f, *p);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int tse_mass_sums(tse_ctx *ctx, const double *C, const double *rho_f,
                  const double *phi, const double *qx, double *out3) {
    if (!out3) return fail(TSE_ERR_ARG, "null argument");
    const int rc = tse_mass_sums_async(ctx, C, rho_f, phi, qx);
    return rc ? rc : tse_reduction_result(ctx, out3, 3);
}

int tse_mass_sums_async(tse_ctx *ctx, const double *C, const double *rho_f,
                        const double *phi, const double *qx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!C || !rho_f || !phi || !qx) return fail(TSE_ERR_ARG, "null argument");
    const Geo &G = ctx->geo;
    const long F = 2L * G.S;
    const int nb = (int)std::min<long>((F + kBlock - 1) / kBlock, 1024);
    if (!ctx->partials || !ctx->red_out)
        return fail(TSE_ERR_STATE, "context scratch not allocated");
    hipLaunchKernelGGL(mass_kernel, dim3(nb), dim3(kBlock), 0, ctx->stream, G,
                       ctx->d_dxyz, C, rho_f, phi, qx, ctx->partials);
    TSE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(mass_reduce_kernel, dim3(1), dim3(kBlock), 0, ctx->stream,
                       ctx->partials, nb, ctx->red_out);
    TSE_HIP_TRY(hipGetLastError());
    TSE_HIP_TRY(hipMemcpyAsync(ctx->red_host, ctx->red_out, 3 * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    return 0;
}

}  // extern "C"
