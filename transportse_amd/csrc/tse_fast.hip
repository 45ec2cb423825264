// tse_fast.hip -- performance forms of the pressure solve (plain FP64 and
// FP32-in-FP64 defect correction).
//
// Operator.  At its fixed point the compute_P update (tse.py:657-697) solves
//   (R + M) P_c - sum_f a_f P_f = gravity_c + R Pn_c + WP_c.
// Scaled by the cell volume V_c the face coefficients become symmetric:
//   W_f = harm(K) / harm(nue) * harm(rho) / ((d1 + d2)/2) * (face area),
// so each cell stores its three "+" faces Wx, Wy, Wz (toward i+1, j+1, k+1),
// the inverse diagonal invD = 1 / (V R + sum of its six W) (0 marks a
// Dirichlet cell) and the right-hand side
//   b = V (R Pn + WP) + Wx+ Gx+ - Wx- Gx- + Wz+ Gz+ - Wz- Gz-
// with G the gravity head of a face (beta_g, tse.py:263-271).  Arrays use the
// natural x-contiguous layout idx = i + nx*(j + ny*k).
//
// Sweep kernel (rb_zmarch).  A block owns a tile of 64 x BY columns and marches
// up a z-chunk, streaming one z-plane per step through LDS (P, Wx, Wy of the
// plane; coalesced 64-wide row loads).  NS full red+black sweeps are fused per
// launch as a phase pipeline: phase h (colour h & 1, red first) updates plane
// k-1-h at step k, after phase h-1 has updated planes k-h .. k-2-h.  A phase
// is valid on the columns at least h+1 cells inside the tile edge, so the
// tile's outer ring of 2*NS columns is halo (recomputed by the neighbouring
// tile) and the inner (64-4NS) x (BY-4NS) columns are written.  A z-chunk
// loads 2*NS extra planes below and above for the same reason.  Input and
// output arrays are distinct (ping-pong), so tiles and chunks never see each
// other's writes within a launch.  Each launch reads every operand once and
// writes the unknown once for NS sweeps.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tse_alloc.h"
#include "tse_fast.h"

namespace tse {

template <typename T, int NS, int BYT, int RY>
__global__ void __launch_bounds__(64 * BYT)
rb_zmarch_kernel(ZGeo g, const T *__restrict__ Pin, T *__restrict__ Pout,
                 const T *__restrict__ Wx, const T *__restrict__ Wy,
                 const T *__restrict__ Wz, const T *__restrict__ invD,
                 const T *__restrict__ rhs, T omega, int zero_init) {
    const int tile = xcd_tile(g);
    if (tile >= g.ntiles) return;  // uniform per block
    rb_zmarch_tile<T, NS, BYT, RY>(g, tile, Pin, Pout, Wx, Wy, Wz, invD, rhs,
                                   omega, zero_init);
}

template <typename T, int NS, int BYT, int RY>
__global__ void __launch_bounds__(64 * BYT)
rb_zmarch_xl_kernel(ZGeo g, const T *__restrict__ Pin, T *__restrict__ Pout,
                    const T *__restrict__ Wx, const T *__restrict__ Wy,
                    const T *__restrict__ Wz, const T *__restrict__ invD,
                    const T *__restrict__ rhs, T omega, int zero_init) {
    const int tile = xcd_tile(g);
    if (tile >= g.ntiles) return;
    rb_zmarch_tile_xl<T, NS, BYT, RY>(g, tile, Pin, Pout, Wx, Wy, Wz, invD, rhs,
                                      omega, zero_init);
}

template <typename T, int NS, int BYT, int RY>
__global__ void __launch_bounds__(64 * BYT)
rb_zmarch_xr_kernel(ZGeo g, const T *__restrict__ Pin, T *__restrict__ Pout,
                    const T *__restrict__ Wx, const T *__restrict__ Wy,
                    const T *__restrict__ Wz, const T *__restrict__ invD,
                    const T *__restrict__ rhs, T omega, int zero_init) {
    const int tile = xcd_tile(g);
    if (tile >= g.ntiles) return;
    rb_zmarch_tile_xr<T, NS, BYT, RY>(g, tile, Pin, Pout, Wx, Wy, Wz, invD, rhs,
                                      omega, zero_init);
}

template <typename T, int NS, int BYT, int RY>
__global__ void __launch_bounds__(64 * BYT)
rb_zmarch_x2_kernel(ZGeo g, const T *__restrict__ Pin, T *__restrict__ Pout,
                    const T *__restrict__ Wx, const T *__restrict__ Wy,
                    const T *__restrict__ Wz, const T *__restrict__ invD,
                    const T *__restrict__ rhs, T omega, int zero_init) {
    static_assert(sizeof(T) == 4 && RY == 2, "variant 3: FP32, two cells per thread");
    const int tile = xcd_tile(g);
    if (tile >= g.ntiles) return;
    rb_zmarch_tile_x2<NS, BYT, true>(g, tile, Pin, Pout, Wx, Wy, Wz, invD, rhs, omega,
                                     zero_init);
}

// Variant 5: cell pairs on the bfloat16 inner operator (rec: 16-byte pair
// records of packed Wx, Wy, Wz and V R; the diagonal summed from the rounded
// values); RB16 (variant 7): the rhs as packed FP16 pairs (scaled).
template <int NS, int BYT, bool RB16>
__global__ void __launch_bounds__(64 * BYT)
rb_zmarch_x5_kernel(ZGeo g, const float *__restrict__ Pin, float *__restrict__ Pout,
                    const uint32_t *__restrict__ rec, const void *__restrict__ rhs,
                    float omega, int zero_init) {
    const int tile = xcd_tile(g);
    if (tile >= g.ntiles) return;
    rb_zmarch_tile_x5<NS, BYT, RB16>(g, tile, Pin, Pout, rec, rhs, omega, zero_init);
}

// ---- operator assembly (from split-layout fields) ---------------------------------------

__device__ inline double f_dsp(const Geo &G, int n, int i, int j, int k) {
    return n == 0 ? G.dx[i] : (n == 1 ? G.dy[j] : G.dz[k]);
}

__device__ inline double f_harm(const Geo &G, const double *H, int i1, int j1,
                                int k1, int i2, int j2, int k2, int n) {
    const double h1 = H[slot(G, i1, j1, k1)], h2 = H[slot(G, i2, j2, k2)];
    if (h1 != 0 && h2 != 0) {
        const double d1 = f_dsp(G, n, i1, j1, k1), d2 = f_dsp(G, n, i2, j2, k2);
        return (d1 + d2) / (d1 / h1 + d2 / h2);
    }
    return 0.0;
}

__device__ inline double f_amean(const Geo &G, const double *X, int i1, int j1,
                                 int k1, int i2, int j2, int k2, int n) {
    const double x1 = X[slot(G, i1, j1, k1)], x2 = X[slot(G, i2, j2, k2)];
    const double d1 = f_dsp(G, n, i1, j1, k1), d2 = f_dsp(G, n, i2, j2, k2);
    return (x1 * d1 + x2 * d2) / (d1 + d2);
}

// Symmetric face transmissibility between (i,j,k) and its + neighbour on
// axis n (caller guarantees the neighbour exists).
__device__ inline double face_W(const Geo &G, const double *K,
                                const double *nue, const double *rho, int i,
                                int j, int k, int n) {
    const int i2 = i + (n == 0), j2 = j + (n == 1), k2 = k + (n == 2);
    const double hk = f_harm(G, K, i2, j2, k2, i, j, k, n);
    const double hn = f_harm(G, nue, i2, j2, k2, i, j, k, n);
    const double hr = f_harm(G, rho, i2, j2, k2, i, j, k, n);
    const double d1 = f_dsp(G, n, i2, j2, k2), d2 = f_dsp(G, n, i, j, k);
    const double area = n == 0 ? G.dy[j] * G.dz[k]
                                : (n == 1 ? G.dx[i] * G.dz[k] : G.dx[i] * G.dy[j]);
    if (hn == 0.0) return 0.0;
    return hk / hn * hr / ((d1 + d2) * 0.5) * area;
}

__device__ inline double face_G(const Geo &G, const double *rho, int i, int j,
                                int k, int n) {
    if (n == 0)
        return f_amean(G, rho, i + 1, j, k, i, j, k, 0) * G.g * 0.5 *
               (G.dx[i + 1] + G.dx[i]) * G.sinb;
    return f_amean(G, rho, i, j, k + 1, i, j, k, 2) * G.g * 0.5 *
           (G.dz[k + 1] + G.dz[k]) * G.cosb;
}

// nxf: row length of the natural layout (nx or nx + 1); the cells of a pad
// column (i = nx) are written inactive.
__global__ void __launch_bounds__(kBlock)
fast_assemble_kernel(Geo G, FastDiv dnx, FastDiv dny, int nxf,
                     const double *__restrict__ Pn, const double *__restrict__ c_f,
                     const double *__restrict__ nue, const double *__restrict__ rho,
                     const double *__restrict__ phi, const double *__restrict__ Kx,
                     const double *__restrict__ Ky, const double *__restrict__ Kz,
                     const double *__restrict__ BP, const double *__restrict__ WP,
                     const double *__restrict__ c_r, double dt, FastOp op) {
    const long N = (long)nxf * G.ny * G.nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const int row = (int)fdiv((uint32_t)t, dnx);
        const int i = (int)t - row * nxf;
        const int k = (int)fdiv((uint32_t)row, dny);
        const int j = row - k * G.ny;
        if (i >= G.nx) {   // pad column
            op.Wx[t] = op.Wy[t] = op.Wz[t] = op.D[t] = op.invD[t] = op.b[t] = 0.0;
            op.Wx32[t] = op.Wy32[t] = op.Wz32[t] = op.invD32[t] = 0.f;
            uint16_t *q = op.rec16 + 8 * (t >> 1) + (t & 1);
            q[0] = q[2] = q[4] = 0;
            q[6] = kBf16Inf;
            continue;
        }
        const int c = slot(G, i, j, k);
        double W[6] = {0, 0, 0, 0, 0, 0};  // x+, x-, y+, y-, z+, z-
        double Gx = 0, Gxm = 0, Gz = 0, Gzm = 0;
        if (i < G.nx - 1) {
            W[0] = face_W(G, Kx, nue, rho, i, j, k, 0);
            Gx = face_G(G, rho, i, j, k, 0);
        }
        if (i > 0) {
            W[1] = face_W(G, Kx, nue, rho, i - 1, j, k, 0);
            Gxm = face_G(G, rho, i - 1, j, k, 0);
        }
        if (j < G.ny - 1) W[2] = face_W(G, Ky, nue, rho, i, j, k, 1);
        if (j > 0) W[3] = face_W(G, Ky, nue, rho, i, j - 1, k, 1);
        if (k < G.nz - 1) {
            W[4] = face_W(G, Kz, nue, rho, i, j, k, 2);
            Gz = face_G(G, rho, i, j, k, 2);
        }
        if (k > 0) {
            W[5] = face_W(G, Kz, nue, rho, i, j, k - 1, 2);
            Gzm = face_G(G, rho, i, j, k - 1, 2);
        }
        const double V = G.dx[i] * G.dy[j] * G.dz[k];
        const double por = phi[c];
        const double R = rho[c] * ((1 - por) * c_r[c] + por * c_f[c]) / dt;
        const double Dg = V * R + (W[0] + W[1]) + (W[2] + W[3]) + (W[4] + W[5]);
        const double bb = V * (R * Pn[c] + WP[c]) + W[0] * Gx - W[1] * Gxm +
                          W[4] * Gz - W[5] * Gzm;
        const bool active = !(BP[c] != 0) && Dg > 0.0;
        const double inv = active ? 1.0 / Dg : 0.0;
        op.Wx[t] = W[0];
        op.Wy[t] = W[2];
        op.Wz[t] = W[4];
        op.D[t] = Dg;
        op.invD[t] = inv;
        op.b[t] = bb;
        op.Wx32[t] = (float)W[0];
        op.Wy32[t] = (float)W[2];
        op.Wz32[t] = (float)W[4];
        op.invD32[t] = (float)inv;
        // variant-5 record: the "+" faces and the storage term rounded to
        // bfloat16 (the sweep sums the diagonal from them), +inf for a cell
        // that does not update
        uint16_t *q = op.rec16 + 8 * (t >> 1) + (t & 1);
        q[0] = bf16_rne(W[0]);
        q[2] = bf16_rne(W[2]);
        q[4] = bf16_rne(W[4]);
        q[6] = active ? bf16_rne(V * R) : kBf16Inf;
    }
}

// ---- layout conversion split <-> natural ------------------------------------------------------

// Natural layout with row length nxf (nx or nx + 1): a pad column reads 0
// and is not written back.
__global__ void __launch_bounds__(kBlock)
split_to_nat_kernel(Geo G, FastDiv dnx, FastDiv dny, int nxf,
                    const double *__restrict__ src, double *__restrict__ dst) {
    const long N = (long)nxf * G.ny * G.nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const int row = (int)fdiv((uint32_t)t, dnx);
        const int i = (int)t - row * nxf;
        const int k = (int)fdiv((uint32_t)row, dny);
        const int j = row - k * G.ny;
        dst[t] = i < G.nx ? src[slot(G, i, j, k)] : 0.0;
    }
}

__global__ void __launch_bounds__(kBlock)
nat_to_split_kernel(Geo G, FastDiv dnx, FastDiv dny, int nxf,
                    const double *__restrict__ src, double *__restrict__ dst) {
    const long N = (long)nxf * G.ny * G.nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const int row = (int)fdiv((uint32_t)t, dnx);
        const int i = (int)t - row * nxf;
        const int k = (int)fdiv((uint32_t)row, dny);
        const int j = row - k * G.ny;
        if (i < G.nx) dst[slot(G, i, j, k)] = src[t];
    }
}

// ---- FP64 residual r = b - A p (and p <- p + e) ------------------------------------------------

__global__ void __launch_bounds__(kBlock)
reduce2_kernel(const double *__restrict__ partials, int n, double *__restrict__ out) {
    double a = 0, c = 0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a += partials[2 * q];
        c += partials[2 * q + 1];
    }
    block_sum2_to(a, c, out);
}

// ---- transport systems (compute_C / compute_T at their fixed point) -------------------------
// The faithful transport records (tse_internal.h kTRec) give, per cell, the
// neighbour coefficients a_f, the constant part X + R*Xn + W - E and the
// diagonal R - omega*M.  At the fixed point of the update (tse.py:870-875,
// 1038-1043) the system is
//     (R - omega M) x_c - sum_f omega a_f x_f = X + R Xn + W - E,
// non-symmetric (upwind), so each cell keeps its own six coefficients
// c_f = omega * a_f (order x+, x-, y+, y-, z+, z-), 1/diagonal and rhs.

__global__ void __launch_bounds__(kBlock)
tr_from_records_kernel(Geo G, FastDiv dnx, FastDiv dny,
                       const double *__restrict__ rec,
                       const uint8_t *__restrict__ mask, double omega, TrOp op) {
    const long N = op.N;
    const long F = G.Fr;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const int row = (int)fdiv((uint32_t)t, dnx);
        const int i = (int)t - row * G.nx;
        const int k = (int)fdiv((uint32_t)row, dny);
        const int j = row - k * G.ny;
        const long c = rec_slot(G, i, j, k);
        const double *r = rec + c;
        const bool active = (mask[c] & 2) != 0;
        const double den = r[10 * F];
        const double inv = (active && den != 0.0) ? 1.0 / den : 0.0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const double cf = omega * r[f * F];
            op.c[f * op.ld + t] = cf;
            op.c32[f * op.ld + t] = (float)cf;
        }
        op.D[t] = den;
        op.invD[t] = inv;
        op.invD32[t] = (float)inv;
        op.b[t] = r[6 * F] + r[7 * F] + r[8 * F] - r[9 * F];
    }
}

template <typename T, int NS, int BY, int VAR>
__global__ void __launch_bounds__(64 * BY)
tr_zmarch_kernel(ZGeo g, const T *__restrict__ Pin, T *__restrict__ Pout,
                 const T *__restrict__ coef, long ld, const T *__restrict__ invD,
                 const T *__restrict__ rhs, T omega, int zero_init) {
    const int tile = xcd_tile(g);
    if (tile >= g.ntiles) return;
    if constexpr (VAR == 0)
        tr_zmarch_tile<T, NS, BY>(g, tile, Pin, Pout, coef, ld, invD, rhs, omega,
                                  zero_init);
    else
        tr_zmarch_tile_xr<T, NS, BY>(g, tile, Pin, Pout, coef, ld, invD, rhs, omega,
                                     zero_init);
}

// Residual as a z-march: a block (kBlock = 64 x 4 threads) takes a tile of
// 64 x 4 columns over one z-chunk; each thread keeps the unknown of planes
// k-1 and k+1 (and, for pressure, Wz of plane k-1) in registers, so each
// operand is read from memory once, and the x / y neighbours, read by the
// adjacent threads in the same plane, come from cache.  The per-cell
// arithmetic is that of p_residual_cell / tr_residual_cell (same terms in
// the same order); only the grouping of the norm partial sums differs.
struct RzGeo {
    int tiles_x, tiles_y, zchunk, ntiles;
};

// R16: the correction's rhs is written as FP16 for sweep variant 7, as FP32
// otherwise.  The FP16 value is r * rscale (rscale a power of two, so the
// product is exact) rounded to FP32 and then to FP16 stochastically, with a
// dither fixed by the cell index: the 13 discarded mantissa bits plus
// hash(t) mod 2^13 carry into the kept ones, and the result is clamped to
// +-65504.  Nearest rounding of a smooth residual leaves a smooth
// (low-frequency) error, the kind SOR removes slowest in the next cycle; the
// dithered error is white.  In MODE 1 the correction read back is e * escale
// (escale = 1 / the rscale of the rhs the sweeps solved with).  hash: the
// lowbias32 integer mix (C. Wellons).
__device__ __forceinline__ uint32_t rhs_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
template <int KIND, int MODE, bool R16>
__global__ void __launch_bounds__(kBlock)
residual_zm_kernel(int nx, int ny, int nz, RzGeo rg, const double *__restrict__ pin,
                   const float *__restrict__ e, double *__restrict__ pout, FastOp pop,
                   TrOp trop, void *__restrict__ rout, double *__restrict__ partials,
                   double escale, double rscale) {
    const long plane = (long)nx * ny;
    const long N = plane * nz;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    double sr = 0.0, sb = 0.0;
    auto X = [&](long o) -> double {
        if constexpr (MODE == 1 && R16) return pin[o] + (double)e[o] * escale;
        return MODE == 1 ? pin[o] + (double)e[o] : pin[o];
    };
    // x / y neighbours of the edge cells are read at clamped indices and
    // discarded by the selects below, so every load of a cell is issued
    // without a branch
    for (int tile = blockIdx.x; tile < rg.ntiles; tile += gridDim.x) {
        const int txy = tile % (rg.tiles_x * rg.tiles_y);
        const int tz = tile / (rg.tiles_x * rg.tiles_y);
        const int i = (txy % rg.tiles_x) * 64 + lx;
        const int j = (txy / rg.tiles_x) * 4 + ly;
        if (i >= nx || j >= ny) continue;
        const bool ie = i < nx - 1, iw = i > 0, jn = j < ny - 1, js = j > 0;
        const long dxe = ie ? 1 : 0, dxw = iw ? 1 : 0, dyn = jn ? nx : 0, dys = js ? nx : 0;
        const int k0 = tz * rg.zchunk, k1 = min(k0 + rg.zchunk, nz);
        long t = (long)i + (long)nx * j + plane * k0;
        double pm = k0 > 0 ? X(t - plane) : 0.0, wzm = 0.0;
        if constexpr (KIND == SYS_P)
            wzm = k0 > 0 ? pop.Wz[t - plane] : 0.0;
        double pc = X(t);
#pragma unroll 2
        for (int k = k0; k < k1; ++k, t += plane) {
            const bool kt = k < nz - 1, kb = k > 0;
            const double pp = X(min(t + plane, N - 1));
            const double xe = X(t + dxe), xw = X(t - dxw), xn = X(t + dyn), xs = X(t - dys);
            if (MODE == 1) pout[t] = pc;
            double r, bb, inv;
            if constexpr (KIND == SYS_P) {
                const double wz = pop.Wz[t], we = pop.Wx[t], ww = pop.Wx[t - dxw];
                const double wn = pop.Wy[t], ws = pop.Wy[t - dys];
                inv = pop.invD[t] > 0.0 ? 1.0 : 0.0;
                bb = pop.b[t];
                double acc = bb;
                acc = ie ? acc + we * xe : acc;
                acc = iw ? acc + ww * xw : acc;
                acc = jn ? acc + wn * xn : acc;
                acc = js ? acc + ws * xs : acc;
                acc = kt ? acc + wz * pp : acc;
                acc = kb ? acc + wzm * pm : acc;
                r = acc - pop.D[t] * pc;
                wzm = wz;
            } else {
                const long Nc = trop.ld;
                const double c0 = trop.c[0 * Nc + t], c1 = trop.c[1 * Nc + t];
                const double c2 = trop.c[2 * Nc + t], c3 = trop.c[3 * Nc + t];
                const double c4 = trop.c[4 * Nc + t], c5 = trop.c[5 * Nc + t];
                inv = trop.invD[t] != 0.0 ? 1.0 : 0.0;
                bb = trop.b[t];
                double acc = bb;
                acc = ie ? acc + c0 * xe : acc;
                acc = iw ? acc + c1 * xw : acc;
                acc = jn ? acc + c2 * xn : acc;
                acc = js ? acc + c3 * xs : acc;
                acc = kt ? acc + c4 * pp : acc;
                acc = kb ? acc + c5 * pm : acc;
                r = acc - trop.D[t] * pc;
            }
            r = inv != 0.0 ? r : 0.0;          // Dirichlet cells: r = 0
            sr += r * r;
            sb += inv != 0.0 ? bb * bb : 0.0;
            if (MODE != 0) {
                if constexpr (R16) {
                    const uint32_t u = __float_as_uint((float)(r * rscale)) +
                                       (rhs_hash((uint32_t)t) & 0x1fffu);
                    const float x = fminf(fmaxf(__uint_as_float(u & ~0x1fffu), -65504.f),
                                          65504.f);
                    ((_Float16 *)rout)[t] = (_Float16)x;
                } else {
                    ((float *)rout)[t] = (float)r;
                }
            }
            pm = pc;
            pc = kt ? pp : 0.0;
        }
    }
    block_sum2_to(sr, sb, partials + 2 * blockIdx.x);
}

static RzGeo make_rzgeo(const Geo &G) {
    RzGeo r;
    r.tiles_x = (G.nx + 63) / 64;
    r.tiles_y = (G.ny + 3) / 4;
    const long txy = (long)r.tiles_x * r.tiles_y;
    // about 4096 tiles (the partial-sum capacity), chunks of at least 8 planes
    r.zchunk = (int)std::min<long>(G.nz, std::max<long>(8, (G.nz * txy + 4095) / 4096));
    r.ntiles = (int)(txy * ((G.nz + r.zchunk - 1) / r.zchunk));
    return r;
}

// ---- host side ---------------------------------------------------------------------------------------

FastSys *system_of(tse_ctx *ctx, int kind) {
    void *&slot_ = kind == SYS_P ? ctx->fast : ctx->fast_tr;
    if (!slot_) {
        FastSys *s = new FastSys();
        s->kind = kind;
        slot_ = s;
    }
    return (FastSys *)slot_;
}

static void free_sys(void *&ptr) {
    FastSys *s = (FastSys *)ptr;
    if (!s) return;
    if (s->buf64) (void)hipFree(s->buf64);
    if (s->buf32) (void)hipFree(s->buf32);
    if (s->partials) (void)hipFree(s->partials);
    if (s->red) (void)hipFree(s->red);
    if (s->red_host) (void)hipHostFree(s->red_host);
    delete s;
    ptr = nullptr;
}

void fast_free(tse_ctx *ctx) {
    free_sys(ctx->fast);
    free_sys(ctx->fast_tr);
}

int ensure_sys(tse_ctx *ctx, FastSys *s) {
    const Geo &G = ctx->geo;
    const int nxf = fast_nx(G, s->kind);
    const long N = (long)nxf * G.ny * G.nz;
    if (!s->partials) {
        if (hipMalloc((void **)&s->partials, sizeof(double) * 2 * 4096) != hipSuccess ||
            hipMalloc((void **)&s->red, sizeof(double) * 2) != hipSuccess ||
            hipHostMalloc((void **)&s->red_host, sizeof(double) * 2) != hipSuccess)
            return fail(TSE_ERR_NOMEM, "fast-form scratch allocation failed");
    }
    if (s->N == N && s->nxf == nxf && s->buf64 && s->buf32) return 0;
    if (s->buf64 || s->buf32) {
        TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (s->buf64) TSE_HIP_TRY(hipFree(s->buf64));
        if (s->buf32) TSE_HIP_TRY(hipFree(s->buf32));
    }
    // "not allocated" until both arrays exist: a failed allocation leaves
    // N = 0 and both pointers null, so a retry allocates again
    s->buf64 = nullptr;
    s->buf32 = nullptr;
    s->N = 0;
    s->nxf = 0;
    s->ld = 0;
    s->ready = false;
    // the arrays are 1088 elements apart beyond a multiple of 64 (FP32:
    // 4.25 KiB, FP64: 8.5 KiB), so that equal offsets in different arrays
    // do not fall on the same HBM channel when N is a large power of two
    // (the 512x512x256 FP32 pressure sweep: 604 -> 577 us per launch)
    const long ld = ((N + 63) / 64) * 64 + 1088;
    void *b64 = nullptr, *b32 = nullptr;
    const bool ok = alloc_both(
        &b64, sizeof(double) * (s->n64() + 2) * ld, &b32,
        sizeof(float) * (s->n32() + 2) * ld,
        [](void **p, size_t n) {
            if (hipMalloc(p, n) == hipSuccess) return true;
            (void)hipGetLastError();
            *p = nullptr;
            return false;
        },
        [](void *p) { (void)hipFree(p); });
    if (!ok) return fail(TSE_ERR_NOMEM, "fast-form operator allocation failed");
    s->buf64 = (double *)b64;
    s->buf32 = (float *)b32;
    s->ld = ld;
    s->N = N;
    s->nxf = nxf;
    return 0;
}

static int grid_n(long n) {
    long b = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min(b, 4096L));
}

// Launches a residual pass and returns ||r||, ||b|| (synchronises).
// r16: write the rhs as FP16 scaled by rscale (pressure systems, sweep
// variant 7) into the FP32 rhs array's memory, and read the correction as
// e * escale.
template <int MODE>
static int residual(tse_ctx *ctx, FastSys *s, const double *pin, const float *e,
                    double *pout, double *rn, double *bn, bool r16 = false,
                    double escale = 1.0, double rscale = 1.0) {
    const Geo G = fast_tiling_geo(ctx->geo, s->kind);
    const RzGeo rg = make_rzgeo(G);
    const int nb = std::max(1, std::min(rg.ntiles, 4096));
    const FastOp pop = s->kind == SYS_P ? s->pop() : FastOp{};
    const TrOp trop = s->kind == SYS_TR ? s->trop() : TrOp{};
    if (s->kind == SYS_P && r16)
        hipLaunchKernelGGL((residual_zm_kernel<SYS_P, MODE, true>), dim3(nb), dim3(kBlock),
                           0, ctx->stream, G.nx, G.ny, G.nz, rg, pin, e, pout, pop, trop,
                           (void *)s->r32(), s->partials, escale, rscale);
    else if (s->kind == SYS_P)
        hipLaunchKernelGGL((residual_zm_kernel<SYS_P, MODE, false>), dim3(nb), dim3(kBlock),
                           0, ctx->stream, G.nx, G.ny, G.nz, rg, pin, e, pout, pop, trop,
                           (void *)s->r32(), s->partials, escale, rscale);
    else
        hipLaunchKernelGGL((residual_zm_kernel<SYS_TR, MODE, false>), dim3(nb), dim3(kBlock),
                           0, ctx->stream, G.nx, G.ny, G.nz, rg, pin, e, pout, pop, trop,
                           (void *)s->r32(), s->partials, escale, rscale);
    TSE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce2_kernel, dim3(1), dim3(kBlock), 0, ctx->stream,
                       s->partials, nb, s->red);
    TSE_HIP_TRY(hipGetLastError());
    TSE_HIP_TRY(hipMemcpyAsync(s->red_host, s->red, 2 * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *rn = std::sqrt(s->red_host[0]);
    *bn = std::sqrt(s->red_host[1]);
    return 0;
}

template <typename T>
static int launch_zmarch(hipStream_t st, const ZGeo &zg, const tse_sweep_plan &pl,
                         const T *Pin, T *Pout, const T *Wx, const T *Wy,
                         const T *Wz, const T *invD, const T *rhs, T omega,
                         int zero_init) {
    const dim3 grid(zg.per_xcd * 8);
    const int ns = pl.ns, by = pl.by, ry = pl.ry, v = pl.variant;
#define TSE_ZM(V_, K_, NS_, BYT_, RY_)                                            \
    if (v == V_ && ns == NS_ && by == BYT_ * RY_ && ry == RY_) {                  \
        hipLaunchKernelGGL((K_<T, NS_, BYT_, RY_>), grid, dim3(64, BYT_), 0, st,   \
                           zg, Pin, Pout, Wx, Wy, Wz, invD, rhs, omega,           \
                           zero_init);                                            \
        TSE_HIP_TRY(hipGetLastError());                                           \
        return 0;                                                                 \
    }
    TSE_ZM(0, rb_zmarch_kernel, 1, 8, 1)
    TSE_ZM(0, rb_zmarch_kernel, 1, 16, 1)
    TSE_ZM(0, rb_zmarch_kernel, 2, 16, 1)
    TSE_ZM(0, rb_zmarch_kernel, 1, 12, 2)
    TSE_ZM(1, rb_zmarch_xl_kernel, 1, 16, 1)
    TSE_ZM(1, rb_zmarch_xl_kernel, 2, 16, 1)
    TSE_ZM(1, rb_zmarch_xl_kernel, 1, 12, 2)
    TSE_ZM(2, rb_zmarch_xr_kernel, 1, 16, 1)
    TSE_ZM(2, rb_zmarch_xr_kernel, 2, 16, 1)
    TSE_ZM(2, rb_zmarch_xr_kernel, 1, 12, 2)
    if constexpr (sizeof(T) == 4) {
        TSE_ZM(0, rb_zmarch_kernel, 1, 16, 2)
        TSE_ZM(0, rb_zmarch_kernel, 2, 16, 2)
        TSE_ZM(0, rb_zmarch_kernel, 2, 8, 4)
        TSE_ZM(0, rb_zmarch_kernel, 3, 16, 1)
        TSE_ZM(1, rb_zmarch_xl_kernel, 2, 12, 2)
        TSE_ZM(1, rb_zmarch_xl_kernel, 2, 16, 2)
        TSE_ZM(1, rb_zmarch_xl_kernel, 3, 16, 1)
        TSE_ZM(1, rb_zmarch_xl_kernel, 3, 12, 2)
        TSE_ZM(2, rb_zmarch_xr_kernel, 2, 12, 2)
        TSE_ZM(2, rb_zmarch_xr_kernel, 2, 16, 2)
        TSE_ZM(2, rb_zmarch_xr_kernel, 3, 16, 1)
        TSE_ZM(2, rb_zmarch_xr_kernel, 3, 12, 2)
        TSE_ZM(2, rb_zmarch_xr_kernel, 3, 16, 2)
        TSE_ZM(3, rb_zmarch_x2_kernel, 2, 16, 2)
        TSE_ZM(3, rb_zmarch_x2_kernel, 3, 16, 2)
    }
#undef TSE_ZM
    return fail(TSE_ERR_ARG, "unsupported sweep plan (ns, by, ry, variant)");
}

// Variants 5 and 7; rhs: FP32 (5) or scaled FP16 (7) values.
static int launch_zmarch5(hipStream_t st, const ZGeo &zg, const tse_sweep_plan &pl,
                          const float *Pin, float *Pout, const FastOp &op,
                          const void *rhs, float omega, int zero_init) {
    const dim3 grid(zg.per_xcd * 8);
    if ((pl.variant == 5 || pl.variant == 7) && pl.by == 32 && pl.ry == 2 &&
        (pl.ns == 2 || pl.ns == 3)) {
        const uint32_t *rec = (const uint32_t *)op.rec16;
        auto k = pl.variant == 7
                     ? (pl.ns == 2 ? rb_zmarch_x5_kernel<2, 16, true>
                                   : rb_zmarch_x5_kernel<3, 16, true>)
                     : (pl.ns == 2 ? rb_zmarch_x5_kernel<2, 16, false>
                                   : rb_zmarch_x5_kernel<3, 16, false>);
        hipLaunchKernelGGL(k, grid, dim3(64, 16), 0, st, zg, Pin, Pout, rec, rhs, omega,
                           zero_init);
        TSE_HIP_TRY(hipGetLastError());
        return 0;
    }
    return fail(TSE_ERR_ARG, "unsupported sweep plan (ns, by, ry, variant)");
}

template <typename T>
static int launch_tr_zmarch(hipStream_t st, const ZGeo &zg, const tse_sweep_plan &pl,
                            const T *Pin, T *Pout, const T *coef, long ld,
                            const T *invD, const T *rhs, T omega, int zero_init) {
    const dim3 grid(zg.per_xcd * 8);
    const int ns = pl.ns, by = pl.by, v = pl.variant;
#define TSE_TZ(V_, NS_, BY_)                                                      \
    if (v == V_ && ns == NS_ && by == BY_) {                                      \
        hipLaunchKernelGGL((tr_zmarch_kernel<T, NS_, BY_, V_>), grid,             \
                           dim3(64, BY_), 0, st, zg, Pin, Pout, coef, ld, invD,   \
                           rhs, omega, zero_init);                                \
        TSE_HIP_TRY(hipGetLastError());                                           \
        return 0;                                                                 \
    }
    TSE_TZ(0, 1, 8)
    TSE_TZ(0, 1, 16)
    TSE_TZ(0, 2, 16)
    TSE_TZ(2, 1, 8)
    TSE_TZ(2, 1, 16)
    TSE_TZ(2, 2, 16)
    if constexpr (sizeof(T) == 4) {
        TSE_TZ(0, 3, 16)
        TSE_TZ(2, 3, 16)
    }
#undef TSE_TZ
    return fail(TSE_ERR_ARG, "unsupported transport sweep plan (ns, by, variant)");
}

bool fast_plan_fits(const tse_sweep_plan &pl, long N, int fp32) {
    if (pl.variant == 5 || pl.variant == 7)
        return N * 8L <= (long)kBufMax;   // 8-byte records per cell
    return pl.variant < 2 || N * (fp32 ? 4L : 8L) <= (long)kBufMax;
}

bool fast_plan_grid_ok(const tse_sweep_plan &pl, int nx) {
    // variants 3, 5 and 7 read cell pairs
    return (pl.variant != 3 && pl.variant != 5 && pl.variant != 7) || nx % 2 == 0;
}

bool fast_plan_ok(const tse_sweep_plan &pl, int fp32, int kind) {
    if (kind == SYS_TR)
        return (pl.variant == 0 || pl.variant == 2) && pl.ry == 1 &&
               ((pl.ns == 1 && (pl.by == 8 || pl.by == 16)) ||
                (pl.ns == 2 && pl.by == 16) || (fp32 && pl.ns == 3 && pl.by == 16));
    // {variant, ns, by, ry}
    static const int both[][4] = {{0, 1, 8, 1},  {0, 1, 16, 1}, {0, 2, 16, 1},
                                  {0, 1, 24, 2}, {1, 1, 16, 1}, {1, 2, 16, 1},
                                  {1, 1, 24, 2}, {2, 1, 16, 1}, {2, 2, 16, 1},
                                  {2, 1, 24, 2}};
    static const int f32[][4] = {{0, 1, 32, 2}, {0, 2, 32, 2}, {0, 2, 32, 4},
                                 {0, 3, 16, 1}, {1, 2, 24, 2}, {1, 2, 32, 2},
                                 {1, 3, 16, 1}, {1, 3, 24, 2}, {2, 2, 24, 2},
                                 {2, 2, 32, 2}, {2, 3, 16, 1}, {2, 3, 24, 2},
                                 {2, 3, 32, 2}, {3, 2, 32, 2}, {3, 3, 32, 2},
                                 {5, 2, 32, 2}, {5, 3, 32, 2}, {7, 2, 32, 2},
                                 {7, 3, 32, 2}};
    auto hit = [&](const int c[4]) {
        return pl.variant == c[0] && pl.ns == c[1] && pl.by == c[2] && pl.ry == c[3];
    };
    for (auto &c : both)
        if (hit(c)) return true;
    if (fp32)
        for (auto &c : f32)
            if (hit(c)) return true;
    return false;
}

ZGeo make_zgeo(const Geo &G, int ns, int by, int zchunk, int edge) {
    ZGeo z{};
    z.nx = G.nx;
    z.ny = G.ny;
    z.nz = G.nz;
    z.edge = edge;
    const int R = 2 * ns, TX = 64 - 2 * R, TY = by - 2 * R;
    z.tiles_x = edge ? span_count(G.nx, 64, R) : (G.nx + TX - 1) / TX;
    z.tiles_y = edge ? span_count(G.ny, by, R) : (G.ny + TY - 1) / TY;
    z.zchunk = std::max(1, std::min(zchunk > 0 ? zchunk : G.nz, G.nz));
    z.nzc = (G.nz + z.zchunk - 1) / z.zchunk;
    z.ntiles = z.tiles_x * z.tiles_y * z.nzc;
    z.per_xcd = (z.ntiles + 7) / 8;
    z.div_tx = make_fastdiv((uint32_t)z.tiles_x);
    z.div_ty = make_fastdiv((uint32_t)z.tiles_y);
    return z;
}

// n FP64 sweeps on the iterate (ping-pong index kept in the system).
static int plain_sweeps(tse_ctx *ctx, FastSys *s, const ZGeo &zg,
                        const tse_sweep_plan &pl, int n, double sor) {
    for (int q = 0; q < n / pl.ns; ++q) {
        int rc;
        if (s->kind == SYS_P) {
            FastOp op = s->pop();
            rc = launch_zmarch<double>(ctx->stream, zg, pl,
                                       s->p(s->cur), s->p(1 - s->cur), op.Wx,
                                       op.Wy, op.Wz, op.invD, op.b, sor, 0);
        } else {
            TrOp op = s->trop();
            rc = launch_tr_zmarch<double>(ctx->stream, zg, pl,
                                          s->p(s->cur), s->p(1 - s->cur), op.c,
                                          op.ld, op.invD, op.b, sor, 0);
        }
        if (rc) return rc;
        s->cur = 1 - s->cur;
    }
    return 0;
}

// n FP32 sweeps on the correction equation A e = r from e = 0; *ecur
// receives the buffer holding e.
static int correction_sweeps(tse_ctx *ctx, FastSys *s, const ZGeo &zg,
                             const tse_sweep_plan &pl, int n, float sor,
                             int *ecur) {
    int e = 0;
    for (int q = 0; q < n / pl.ns; ++q) {
        int rc;
        if (s->kind == SYS_P && (pl.variant == 5 || pl.variant == 7)) {
            FastOp op = s->pop();
            rc = launch_zmarch5(ctx->stream, zg, pl, s->e(e), s->e(1 - e), op,
                                s->r32(), sor, q == 0);
        } else if (s->kind == SYS_P) {
            FastOp op = s->pop();
            rc = launch_zmarch<float>(ctx->stream, zg, pl,
                                      s->e(e), s->e(1 - e), op.Wx32, op.Wy32,
                                      op.Wz32, op.invD32, s->r32(), sor, q == 0);
        } else {
            TrOp op = s->trop();
            rc = launch_tr_zmarch<float>(ctx->stream, zg, pl, s->e(e),
                                         s->e(1 - e), op.c32, op.ld, op.invD32,
                                         s->r32(), sor, q == 0);
        }
        if (rc) return rc;
        e = 1 - e;
    }
    *ecur = e;
    return 0;
}

static int load_iterate(tse_ctx *ctx, FastSys *s, const double *X_split) {
    const Geo &G = ctx->geo;
    s->cur = 0;
    hipLaunchKernelGGL(split_to_nat_kernel, dim3(grid_n(s->N)), dim3(kBlock), 0,
                       ctx->stream, G, make_fastdiv(s->nxf), make_fastdiv(G.ny), s->nxf,
                       X_split, s->p(0));
    TSE_HIP_TRY(hipGetLastError());
    s->ready = true;
    return 0;
}

static int sys_result(tse_ctx *ctx, FastSys *s, double *X_split) {
    if (!s->ready || !X_split) return fail(TSE_ERR_STATE, "fast form not set up");
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(nat_to_split_kernel, dim3(grid_n(s->N)), dim3(kBlock), 0,
                       ctx->stream, G, make_fastdiv(s->nxf), make_fastdiv(G.ny), s->nxf,
                       s->p(s->cur), X_split);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

static int sys_residual(tse_ctx *ctx, FastSys *s, double *rel) {
    if (!s->ready || !rel) return fail(TSE_ERR_STATE, "fast form not set up");
    double rn, bn;
    int rc = residual<0>(ctx, s, s->p(s->cur), nullptr, nullptr, &rn, &bn);
    if (rc) return rc;
    *rel = rn / (bn > 0 ? bn : 1.0);
    return 0;
}

static int sys_sweeps(tse_ctx *ctx, FastSys *s, const tse_sweep_plan *plan,
                      int n_sweeps, double sor, int fp32) {
    if (!s->ready || !plan) return fail(TSE_ERR_STATE, "fast form not set up");
    if (!fast_plan_ok(*plan, fp32, s->kind) || n_sweeps < 0 || n_sweeps % plan->ns)
        return fail(TSE_ERR_ARG, "unsupported plan or n_sweeps not a multiple of ns");
    if (!fast_plan_fits(*plan, s->N, fp32))
        return fail(TSE_ERR_ARG, "sweep variants 2, 3, 5 and 7 need fields of at most 1 GiB "
                                 "(variants 5 and 7: 8 bytes per cell)");
    if (!fast_plan_grid_ok(*plan, s->nxf))
        return fail(TSE_ERR_ARG, "sweep variants 3, 5 and 7 need an even nx");
    const ZGeo zg = make_zgeo(fast_tiling_geo(ctx->geo, s->kind), plan->ns, plan->by,
                              plan->zchunk, plan->variant >= 2);
    if (!fp32) return plain_sweeps(ctx, s, zg, *plan, n_sweeps, sor);
    int e;  // FP32 timing path: correction sweeps on the last residual
    return correction_sweeps(ctx, s, zg, *plan, n_sweeps, (float)sor, &e);
}

// Young's relation (see tse_sor_model): 1 and the optimum when the
// contraction per sweep lam at omega shows omega below it, 2 when it does
// but the optimum is not resolved (mu >= 1 or past the cap: *omega_opt =
// the cap), else 0.
static int sor_model(double omega, double lam, double *omega_opt) {
    if (std::log(lam) <= 0.9 * std::log(omega - 1.0)) return 0;   // about omega - 1
    const double mu = (lam + omega - 1.0) / (omega * std::sqrt(lam));
    const double cap = 2.0 - 0x1p-10;
    const double opt = mu >= 1.0 ? cap : 2.0 / (1.0 + std::sqrt(1.0 - mu * mu));
    *omega_opt = std::min(cap, opt);
    return opt >= cap ? 2 : 1;
}

// Whether a cycle of n sweeps at omega resolves a contraction near
// omega - 1: the cycle must span n |log(omega - 1)| >= 1/4 (a reduction by
// e^-1/4 at that rate); shorter cycles (omega near 2) are dominated by the
// transient after their restart from e = 0.
static bool resolves_contraction(int n, double omega) {
    return omega > 1.0 && omega < 2.0 && n * std::fabs(std::log(omega - 1.0)) >= 0.25;
}

static int sys_solve(tse_ctx *ctx, FastSys *s, const tse_fast_params *prm,
                     tse_fast_result *out) {
    if (!s->ready || !prm || !out) return fail(TSE_ERR_STATE, "fast form not set up");
    const int fp32 = prm->form == TSE_FORM_MIXED;
    if (prm->form != TSE_FORM_MIXED && prm->form != TSE_FORM_PLAIN64)
        return fail(TSE_ERR_ARG, "form must be PLAIN64 or MIXED");
    const tse_sweep_plan &pl = prm->plan;
    const int inner = prm->inner;
    int chunk = inner;
    if (!fast_plan_ok(pl, fp32, s->kind) || inner < pl.ns || inner % pl.ns ||
        prm->max_sweeps < 0)
        return fail(TSE_ERR_ARG, "unsupported plan / inner sweep count");
    if (prm->cycle_mode != TSE_CYCLE_FIXED && prm->cycle_mode != TSE_CYCLE_ADAPTIVE)
        return fail(TSE_ERR_ARG, "cycle_mode must be TSE_CYCLE_FIXED or TSE_CYCLE_ADAPTIVE");
    if (prm->adapt_sor != 0 && prm->adapt_sor != 1)
        return fail(TSE_ERR_ARG, "adapt_sor must be 0 or 1");
    if (!fast_plan_fits(pl, s->N, fp32))
        return fail(TSE_ERR_ARG, "sweep variants 2, 3, 5 and 7 need fields of at most 1 GiB "
                                 "(variants 5 and 7: 8 bytes per cell)");
    if (!fast_plan_grid_ok(pl, s->nxf))
        return fail(TSE_ERR_ARG, "sweep variants 3, 5 and 7 need an even nx");
    const ZGeo zg = make_zgeo(fast_tiling_geo(ctx->geo, s->kind), pl.ns, pl.by, pl.zchunk,
                              pl.variant >= 2);
    double rn, bn;
    // variant 7: the correction's rhs in FP16, scaled by a power of two near
    // 1 / RMS of the last residual whose norm is known (for the first one a
    // norm-only pass), so no entry can exceed the FP16 range while the
    // residual shrinks (|r_i| <= sqrt(N) RMS < 65504 for N < 2^32)
    const bool r16 = pl.variant == 7;
    auto rhs_scale = [&](double rnorm) {
        if (!(rnorm > 0) || !std::isfinite(rnorm)) return 1.0;
        int ex;
        (void)std::frexp(std::sqrt((double)s->N) / rnorm, &ex);
        return std::ldexp(1.0, std::min(std::max(ex - 1, -120), 120));
    };
    double rscale = 1.0;
    int rc = 0;
    if (fp32 && r16) {
        rc = residual<0>(ctx, s, s->p(s->cur), nullptr, nullptr, &rn, &bn);
        if (rc) return rc;
        rscale = rhs_scale(rn);
    }
    rc = fp32 ? residual<2>(ctx, s, s->p(s->cur), nullptr, nullptr, &rn, &bn, r16, 1.0,
                            rscale)
              : residual<0>(ctx, s, s->p(s->cur), nullptr, nullptr, &rn, &bn);
    if (rc) return rc;
    const double bs = bn > 0 ? bn : 1.0;
    out->initial_rel = rn / bs;
    out->contraction = 0.0;
    double rel = rn / bs;
    int sweeps = 0, cycles = 0, switches = 0;
    double sor = prm->sor;
    while (rel > prm->tol && sweeps < prm->max_sweeps) {
        // the last cycle ends at max_sweeps (rounded up to the plan's ns)
        const int left = (prm->max_sweeps - sweeps + pl.ns - 1) / pl.ns * pl.ns;
        const int run = std::min(chunk, left);
        if (fp32) {
            int e;
            rc = correction_sweeps(ctx, s, zg, pl, run, (float)sor, &e);
            if (rc) return rc;
            const double escale = 1.0 / rscale;   // exact: a power of two
            rscale = r16 ? rhs_scale(rn) : 1.0;
            rc = residual<1>(ctx, s, s->p(s->cur), s->e(e), s->p(1 - s->cur),
                             &rn, &bn, r16, escale, rscale);
            if (rc) return rc;
            s->cur = 1 - s->cur;
        } else {
            rc = plain_sweeps(ctx, s, zg, pl, run, sor);
            if (rc) return rc;
            rc = residual<0>(ctx, s, s->p(s->cur), nullptr, nullptr, &rn, &bn);
            if (rc) return rc;
        }
        sweeps += run;
        cycles += 1;
        const double rel_prev = rel;
        rel = rn / bs;
        if (!(rel == rel)) break;  // NaN
        // per-sweep contraction of this cycle (the first one starts with the
        // fast-decaying components, so only later cycles are reported)
        if (cycles > 1 && rel < rel_prev && rel > 0 && resolves_contraction(run, sor)) {
            out->contraction = std::exp(std::log(rel / rel_prev) / run);
            // adaptive SOR: a cycle below the optimum moves the factor to
            // the estimate (the next cycle starts from e = 0 anyway)
            double opt;
            if (prm->adapt_sor && switches < 2 && rel > prm->tol &&
                sor_model(sor, out->contraction, &opt) == 1) {
                const double u = std::ceil(-std::log2(2.0 - opt) * 128.0) / 128.0;
                const double next = 2.0 - std::exp2(-u);
                if (next > sor) {
                    sor = next;
                    ++switches;
                }
            }
        }
        if (prm->cycle_mode == TSE_CYCLE_ADAPTIVE) {
            // sweeps still needed at the last cycle's rate, clamped; with
            // adaptive SOR the second cycle stays `inner` long, so the first
            // estimate of the factor comes early
            long next = inner;
            if (rel < rel_prev && rel > prm->tol && rel_prev > 0 &&
                !(prm->adapt_sor && cycles == 1)) {
                const double rate = std::log(rel / rel_prev) / run;     // < 0
                // a variant-5 or -7 cycle cannot reduce the error by much
                // more than its inner operator's 2^-8 accuracy: aim at most
                // two decades per cycle, so its sweeps are not spent past
                // that
                const double target = pl.variant == 5 || pl.variant == 7
                                          ? std::max(prm->tol / rel, 1e-2)
                                          : prm->tol / rel;
                const double need = std::log(target) / rate;             // > 0
                next = (long)std::ceil(need * 1.02);
            }
            next = std::min<long>(std::max<long>(next, inner), 8L * inner);
            next = (next + pl.ns - 1) / pl.ns * pl.ns;
            chunk = (int)next;
        }
    }
    out->sweeps = sweeps;
    out->cycles = cycles;
    out->rel_residual = rel;
    out->converged = rel <= prm->tol ? 1 : 0;
    out->sor = sor;
    return 0;
}

}  // namespace tse

using namespace tse;

extern "C" {

int tse_fast_p_setup(tse_ctx *ctx, const tse_p_fields *f, double dt) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !f->P || !f->c_f || !f->nue || !f->rho_f || !f->phi ||
        !f->Kx || !f->Ky || !f->Kz || !f->BP || !f->WP || !f->c_r)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (!(dt > 0)) return fail(TSE_ERR_ARG, "dt must be > 0");
    FastSys *s = system_of(ctx, SYS_P);
    rc = ensure_sys(ctx, s);
    if (rc) return rc;
    const Geo &G = ctx->geo;
    const double *Pn = f->Pn ? f->Pn : f->P;
    hipLaunchKernelGGL(fast_assemble_kernel, dim3(grid_n(s->N)), dim3(kBlock), 0,
                       ctx->stream, G, make_fastdiv(s->nxf), make_fastdiv(G.ny), s->nxf, Pn,
                       f->c_f, f->nue, f->rho_f, f->phi, f->Kx, f->Ky, f->Kz,
                       f->BP, f->WP, f->c_r, dt, s->pop());
    TSE_HIP_TRY(hipGetLastError());
    return load_iterate(ctx, s, f->P);
}

int tse_fast_p_result(tse_ctx *ctx, double *P) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_result(ctx, system_of(ctx, SYS_P), P);
}

int tse_fast_p_operator(tse_ctx *ctx, double *Wx, double *Wy, double *Wz, double *D,
                        double *b) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    FastSys *s = system_of(ctx, SYS_P);
    if (!s->ready) return fail(TSE_ERR_STATE, "fast form not set up");
    const FastOp op = s->pop();
    const double *src[5] = {op.Wx, op.Wy, op.Wz, op.D, op.b};
    double *dst[5] = {Wx, Wy, Wz, D, b};
    for (int a = 0; a < 5; ++a)
        if (dst[a])
            TSE_HIP_TRY(hipMemcpyAsync(dst[a], src[a], sizeof(double) * s->N,
                                       hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}

int tse_sor_model(double omega, double contraction, double *omega_opt) {
    if (!omega_opt || !(contraction > 0.0 && contraction < 1.0) ||
        !(omega > 1.0 && omega < 2.0))
        return fail(TSE_ERR_ARG, "tse_sor_model: need 0 < contraction < 1 < omega < 2");
    return sor_model(omega, contraction, omega_opt) ? 1 : 0;
}

int tse_fast_p_residual(tse_ctx *ctx, double *rel) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_residual(ctx, system_of(ctx, SYS_P), rel);
}

int tse_fast_p_sweeps(tse_ctx *ctx, const tse_sweep_plan *plan,
                      int32_t n_sweeps, double sor, int32_t fp32) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_sweeps(ctx, system_of(ctx, SYS_P), plan, n_sweeps, sor, fp32);
}

int tse_fast_p_solve(tse_ctx *ctx, const tse_fast_params *prm,
                     tse_fast_result *out) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_solve(ctx, system_of(ctx, SYS_P), prm, out);
}

// Transport: assemble with the faithful record kernels (reference arithmetic,
// Xn = the unknown at entry), convert to the natural-layout operator, load
// the unknown as the iterate.  f->Cn / f->Tn receive the entry copy.
static int tr_setup_common(tse_ctx *ctx, double *X, double omega) {
    FastSys *s = system_of(ctx, SYS_TR);
    int rc = ensure_sys(ctx, s);
    if (rc) return rc;
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(tr_from_records_kernel, dim3(grid_n(s->N)), dim3(kBlock), 0,
                       ctx->stream, G, make_fastdiv(G.nx), make_fastdiv(G.ny),
                       ctx->rec, ctx->mask, omega, s->trop());
    TSE_HIP_TRY(hipGetLastError());
    return load_iterate(ctx, s, X);
}

int tse_fast_c_setup(tse_ctx *ctx, const tse_c_fields *f, double dt,
                     double omega) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !f->C || !f->Cn || !f->D_fx || !f->D_fy || !f->D_fz || !f->phi ||
        !f->rho_f || !f->qx || !f->qy || !f->qz || !f->BC || !f->WC ||
        !f->RC || !f->e || !f->lf_l || !f->lf_t)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (!(dt > 0)) return fail(TSE_ERR_ARG, "dt must be > 0");
    TSE_HIP_TRY(hipMemcpyAsync(f->Cn, f->C, 2ull * ctx->geo.S * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
    rc = launch_c_assemble(ctx, f, dt, omega);
    if (rc) return rc;
    return tr_setup_common(ctx, f->C, omega);
}

int tse_fast_t_setup(tse_ctx *ctx, const tse_t_fields *f, double dt,
                     double omega) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !f->T || !f->Tn || !f->l_f || !f->cp_f || !f->rho_f || !f->phi ||
        !f->qx || !f->qy || !f->qz || !f->BT || !f->WT || !f->e || !f->l_r ||
        !f->cp_r || !f->rho_r || !f->lt_l || !f->lt_t)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (!(dt > 0)) return fail(TSE_ERR_ARG, "dt must be > 0");
    TSE_HIP_TRY(hipMemcpyAsync(f->Tn, f->T, 2ull * ctx->geo.S * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
    rc = launch_t_assemble(ctx, f, dt, omega);
    if (rc) return rc;
    return tr_setup_common(ctx, f->T, omega);
}

int tse_fast_mark_loaded(tse_ctx *ctx, int32_t kind) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (kind != SYS_P && kind != SYS_TR) return fail(TSE_ERR_ARG, "kind must be 0 or 1");
    FastSys *s = (FastSys *)(kind == SYS_P ? ctx->fast : ctx->fast_tr);
    if (!s || !s->buf64) return fail(TSE_ERR_STATE, "fast system never set up");
    s->cur = 0;
    s->ready = true;
    return 0;
}

int tse_fast_current(tse_ctx *ctx, int32_t kind, int32_t *cur) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (kind != SYS_P && kind != SYS_TR) return fail(TSE_ERR_ARG, "kind must be 0 or 1");
    if (!cur) return fail(TSE_ERR_ARG, "null pointer");
    FastSys *s = (FastSys *)(kind == SYS_P ? ctx->fast : ctx->fast_tr);
    if (!s || !s->ready) return fail(TSE_ERR_STATE, "fast form not set up");
    *cur = s->cur;
    return 0;
}

int tse_fast_tr_solve(tse_ctx *ctx, const tse_fast_params *prm,
                      tse_fast_result *out) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_solve(ctx, system_of(ctx, SYS_TR), prm, out);
}

int tse_fast_tr_result(tse_ctx *ctx, double *X) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_result(ctx, system_of(ctx, SYS_TR), X);
}

int tse_fast_tr_residual(tse_ctx *ctx, double *rel) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_residual(ctx, system_of(ctx, SYS_TR), rel);
}

int tse_fast_tr_sweeps(tse_ctx *ctx, const tse_sweep_plan *plan,
                       int32_t n_sweeps, double sor, int32_t fp32) {
    int rc = check_ctx(ctx);
    return rc ? rc : sys_sweeps(ctx, system_of(ctx, SYS_TR), plan, n_sweeps, sor, fp32);
}

}  // extern "C"

// ---- the performance forms behind compute_P / compute_C / compute_T -----------------------
// tse_compute_p / _c / _t with form TSE_FORM_PLAIN64 or TSE_FORM_MIXED: one
// attempt solves the fixed-point system of the reference update at dt to
// the relative residual prm->conv within prm->max_it sweeps; the time-step
// control around the attempts is the reference's (tse.py:724-740 for
// compute_P, 896-914 compute_C, 1065-1079 compute_T):
//  * converged: t += dt, and if the attempt took fewer than max_it/2 sweeps,
//    dt = min(dt * inc_it, max_dt);
//  * not converged within max_it sweeps (or the residual turned NaN, which
//    the reference's norm reports as 1 until max_it): dt /= 2 and the unknown
//    is reset to its value at entry; dt <= min_dt ends the call with
//    converged = 0.
// Transport with omega == 0 is accepted after its first launch, as the
// reference accepts after one iteration.  num_it counts sweeps over all
// attempts; last_diff is the last relative residual.
namespace tse {

// ns sweeps per launch; explicit: one sweep per launch (transport, omega 0)
static tse_sweep_plan default_fast_plan(const tse_ctx *ctx, int kind, int fp32,
                                        bool explicit_step) {
    const Geo &G = ctx->geo;
    const long N = (long)fast_nx(G, kind) * G.ny * G.nz;
    tse_sweep_plan pl{explicit_step ? 1 : 2, 16, 1, 0, 2};
    if (kind == SYS_P && fp32) {
        // the cell-pair sweep on the bfloat16 inner operator with the FP16
        // rhs (pressure rows are even) where its records fit, else the
        // unrolled-ring sweep
        pl = tse_sweep_plan{3, 32, 2, 0, 7};
        if (!fast_plan_fits(pl, N, fp32)) pl.variant = 2;
    }
    if (!fast_plan_fits(pl, N, fp32)) {  // fields over 1 GiB: no buffer loads
        pl = tse_sweep_plan{explicit_step ? 1 : 2, 16, 1, 0, 0};
    }
    return pl;
}

int fast_compute(tse_ctx *ctx, int kind, const void *fields,
                 const tse_sor_params *prm, tse_result *out) {
    const int fp32 = prm->form == TSE_FORM_MIXED;
    if (prm->form != TSE_FORM_MIXED && prm->form != TSE_FORM_PLAIN64)
        return fail(TSE_ERR_ARG, "form must be FAITHFUL64, PLAIN64 or MIXED");
    if (prm->max_it < 1) return fail(TSE_ERR_ARG, "max_it must be >= 1");
    if (!(prm->dt > 0.0)) return fail(TSE_ERR_ARG, "dt must be > 0");
    const int which = kind;  // 0 pressure, 1 compute_C, 2 compute_T
    const int sys_kind = which == 0 ? SYS_P : SYS_TR;
    const Geo &G = ctx->geo;
    const size_t bytes = 2ull * G.S * sizeof(double);
    const tse_p_fields *fp = which == 0 ? (const tse_p_fields *)fields : nullptr;
    const tse_c_fields *fc = which == 1 ? (const tse_c_fields *)fields : nullptr;
    const tse_t_fields *ft = which == 2 ? (const tse_t_fields *)fields : nullptr;
    double *X = fp ? fp->P : (fc ? fc->C : ft->T);
    double *Xn = fp ? fp->Pn : (fc ? fc->Cn : ft->Tn);
    // the unknown at entry (compute_P's Pn, tse.py:656; the transport setups
    // copy it themselves)
    TSE_HIP_TRY(hipMemcpyAsync(Xn, X, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    const bool explicit_step = which != 0 && prm->omega == 0.0;
    // omega == 0: one red+black sweep, as the reference's single iteration
    const tse_sweep_plan pl = default_fast_plan(ctx, sys_kind, fp32, explicit_step);
    tse_fast_params fpm{};
    fpm.form = prm->form;
    fpm.plan = pl;
    // first cycle about max_it/8 sweeps (at most 64), then sized from the
    // observed rate (adaptive cycles), all multiples of the plan's ns
    const int first = std::max(1, std::min(64, prm->max_it / 8));
    fpm.inner = std::max(pl.ns, first / pl.ns * pl.ns);
    fpm.cycle_mode = TSE_CYCLE_ADAPTIVE;
    fpm.sor = prm->sor;
    fpm.tol = explicit_step ? 0.0 : prm->conv;
    fpm.max_sweeps = explicit_step ? pl.ns : prm->max_it;
    double dt = prm->dt, t = 0.0, rel = 0.0;
    int total = 0, halvings = 0, status = FAILED;
    FastSys *s = system_of(ctx, sys_kind);
    for (;;) {
        int rc;
        if (fp) {
            rc = tse_fast_p_setup(ctx, fp, dt);
        } else if (fc) {
            rc = tse_fast_c_setup(ctx, fc, dt, prm->omega);
        } else {
            rc = tse_fast_t_setup(ctx, ft, dt, prm->omega);
        }
        if (rc) return rc;
        tse_fast_result r{};
        rc = sys_solve(ctx, s, &fpm, &r);
        if (rc) return rc;
        total += r.sweeps;
        rel = r.rel_residual;
        if (r.converged || explicit_step) {
            rc = sys_result(ctx, s, X);
            if (rc) return rc;
            t += dt;
            if ((double)r.sweeps < prm->max_it / 2.0) dt = std::min(dt * prm->inc_it, prm->max_dt);
            status = CONVERGED;
            break;
        }
        dt = dt / 2;  // tse.py:733-738: halve, reset the unknown
        halvings += 1;
        TSE_HIP_TRY(hipMemcpyAsync(X, Xn, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        if (dt <= prm->min_dt) {
            status = FAILED;
            break;
        }
    }
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    out->t = t;
    out->dt = dt;
    out->converged = status == CONVERGED ? 1 : 0;
    out->num_it = total;
    out->num_halvings = halvings;
    out->status = status;
    out->last_diff = rel;
    return 0;
}

}  // namespace tse
