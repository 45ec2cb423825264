// tse_step.hip -- device-resident pieces of the SIA step around the solvers:
// the closed-form linear EOS / conductivity update of the Picard loop
// (tse.py:1422-1423 calls config.fluid_properties) and FP64 mass-balance
// sums (tse.py:1461-1463 calls config.mass_balance).  Both act on
// split-layout fields of one context.
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

}  // namespace tse

using namespace tse;

extern "C" {

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
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!C || !rho_f || !phi || !qx || !out3) return fail(TSE_ERR_ARG, "null argument");
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
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < 3; ++c) out3[c] = ctx->red_host[c];
    return 0;
}

}  // extern "C"
