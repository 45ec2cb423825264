// tse_fast.h -- shared pieces of the performance forms: z-march tile
// functions, per-cell residuals and the per-context system state, used by
// the single-system entry points (tse_fast.hip) and the batched engine
// (tse_batch.hip).
#pragma once

#include <cstdint>

#include "tse_internal.h"

namespace tse {

struct ZGeo {
    int nx, ny, nz;
    int tiles_x, tiles_y, zchunk, nzc;
    int ntiles, per_xcd;
    FastDiv div_tx, div_ty;
    int edge;   // boundary-aware tiling (span_of), used by the variant-2 kernels
};

// Boundary-aware tiling of an extent n by tiles W cells wide with a halo of
// R on interior edges: a tile edge on the grid boundary needs no halo (there
// is no neighbour to lag behind).  One tile if n <= W; else the first tile
// starts at 0 and owns [0, W-R), middle tiles own W-2R cells each behind an
// R halo, and the last tile ends at n and owns the rest (at most W-R cells,
// so its left halo is at least R).
__host__ __device__ inline int span_count(int n, int W, int R) {
    if (n <= W) return 1;
    const int T = W - 2 * R, rest = n - 2 * (W - R);
    return 2 + (rest > 0 ? (rest + T - 1) / T : 0);
}

struct Span {
    int x0, lo, hi;      // first cell of the tile; owned cells [lo, hi)
    bool lfree, hfree;   // tile edge on the grid boundary
};

__device__ inline Span span_of(int t, int tiles, int n, int W, int R) {
    Span s;
    if (tiles == 1) {
        s.x0 = 0;
        s.lo = 0;
        s.hi = n;
        s.lfree = s.hfree = true;
        return s;
    }
    const int T = W - 2 * R;
    s.lo = t == 0 ? 0 : (W - R) + (t - 1) * T;
    s.hi = t == tiles - 1 ? n : (W - R) + t * T;
    s.x0 = t == 0 ? 0 : (t == tiles - 1 ? n - W : s.lo - R);
    s.lfree = t == 0;
    s.hfree = t == tiles - 1;
    return s;
}

// Span of tile t: boundary-aware (edge) or the plain tiling, where every
// tile carries an R halo on both sides and owns W-2R cells.
__device__ inline Span tile_span(bool edge, int t, int tiles, int n, int W, int R) {
    if (edge) return span_of(t, tiles, n, W, R);
    Span s;
    const int T = W - 2 * R;
    s.lo = t * T;
    s.hi = min(s.lo + T, n);
    s.x0 = s.lo - R;
    s.lfree = s.hfree = false;
    return s;
}
constexpr int kFar = 1 << 20;   // distance to a tile edge on the grid boundary

// Blocks are dispatched round-robin over the 8 XCDs; map block b to tile
// (b % 8) * per_xcd + b / 8 so each XCD (own L2) works on a contiguous band of
// tiles and shares their halos.
__device__ inline int xcd_tile(const ZGeo &g) {
    const int b = blockIdx.x;
    return (b & 7) * g.per_xcd + (b >> 3);
}

// One tile of the fused red/black z-march (blocks of 64 x BYT threads).
// A block that processes several tiles in turn must barrier between them.
template <typename T, int NS, int BYT, int RY>
__device__ __forceinline__ void
rb_zmarch_tile(const ZGeo &g, int tile, const T *__restrict__ Pin,
               T *__restrict__ Pout, const T *__restrict__ Wx,
               const T *__restrict__ Wy, const T *__restrict__ Wz,
               const T *__restrict__ invD, const T *__restrict__ rhs, T omega,
               int zero_init) {
    constexpr int R = 2 * NS;          // halo ring width / number of phases
    constexpr int BX = 64;
    constexpr int BY = BYT * RY;       // tile rows; thread rows ty + q*BYT
    constexpr int TX = BX - 2 * R;
    constexpr int TY = BY - 2 * R;
    // LDS ring of z-planes: the plane stored at step k and the R planes in
    // phases (k-1 .. k-R), plus one spare so the slot overwritten at step k
    // was last read at step k-2.
    constexpr int NSLOT = R + 2;
    static_assert(TY >= 1, "tile too small");
    __shared__ T sP[NSLOT][BY][BX];
    __shared__ T sWx[NSLOT][BY][BX];
    __shared__ T sWy[NSLOT][BY][BX];

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);   // tile / tiles_x
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);    // / tiles_y
    const int ty = zc - zc2 * g.tiles_y;
    const int lx = threadIdx.x;
    const int i = tx * TX - R + lx;
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const long plane = (long)g.nx * g.ny;
    const int dxe = min(lx, BX - 1 - lx);

    int ly[RY], d[RY], par[RY];
    bool in_dom[RY];
    long col[RY];
    T p[RY][R + 2], rinv[RY][R + 1], rb[RY][R + 1], rwz[RY][R + 2];
    T np[RY], ni[RY], nb[RY], nwz[RY], nwx[RY], nwy[RY];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        ly[q] = threadIdx.y + q * BYT;
        const int j = ty * TY - R + ly[q];
        in_dom[q] = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
        d[q] = min(dxe, min(ly[q], BY - 1 - ly[q]));
        col[q] = in_dom[q] ? (long)i + (long)g.nx * j : 0;
        par[q] = (i + j) & 1;
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
            p[q][r] = T(0);
            rwz[q][r] = T(0);
        }
#pragma unroll
        for (int r = 0; r < R + 1; ++r) {
            rinv[q][r] = T(0);
            rb[q][r] = T(0);
        }
        // operands of the next plane, loaded one step ahead
        np[q] = ni[q] = nb[q] = nwz[q] = nwx[q] = nwy[q] = T(0);
        if (in_dom[q] && kl0 <= kl1) {
            const long o = col[q] + plane * kl0;
            np[q] = zero_init ? T(0) : Pin[o];
            ni[q] = invD[o];
            nb[q] = rhs[o];
            nwz[q] = Wz[o];
            nwx[q] = Wx[o];
            nwy[q] = Wy[o];
        }
    }

    int s0 = kl0 % NSLOT;
    for (int k = kl0; k <= k1 + R - 1; ++k) {
#pragma unroll
        for (int q = 0; q < RY; ++q) {
            const T p0 = np[q], i0 = ni[q], b0 = nb[q], wz0 = nwz[q];
            const T wx0 = nwx[q], wy0 = nwy[q];
            np[q] = ni[q] = nb[q] = nwz[q] = nwx[q] = nwy[q] = T(0);
            if (in_dom[q] && k + 1 <= kl1) {
                const long o = col[q] + plane * (k + 1);
                np[q] = zero_init ? T(0) : Pin[o];
                ni[q] = invD[o];
                nb[q] = rhs[o];
                nwz[q] = Wz[o];
                nwx[q] = Wx[o];
                nwy[q] = Wy[o];
            }
#pragma unroll
            for (int r = R + 1; r > 0; --r) {
                p[q][r] = p[q][r - 1];
                rwz[q][r] = rwz[q][r - 1];
            }
#pragma unroll
            for (int r = R; r > 0; --r) {
                rinv[q][r] = rinv[q][r - 1];
                rb[q][r] = rb[q][r - 1];
            }
            p[q][0] = p0;
            rwz[q][0] = wz0;
            rinv[q][0] = i0;
            rb[q][0] = b0;
            sP[s0][ly[q]][lx] = p0;
            sWx[s0][ly[q]][lx] = wx0;
            sWy[s0][ly[q]][lx] = wy0;
        }
        // One barrier per step: phase h reads in-plane values of plane
        // k-1-h written by phase h-1 during step k-1 (or stored at step
        // k-1 for h = 0) and writes only its own colour; its z neighbours
        // are this thread's registers.
        __syncthreads();
#pragma unroll
        for (int h = 0; h < R; ++h) {
            const int m = k - 1 - h;
            int s = s0 - 1 - h;
            if (s < 0) s += NSLOT;
            const bool mok = m >= kl0 && m <= kl1;
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                if (mok && d[q] >= h + 1 && in_dom[q] &&
                    ((par[q] + m) & 1) == (h & 1) && rinv[q][1 + h] > T(0)) {
                    const int y = ly[q];
                    T acc = rb[q][1 + h];
                    acc += sWx[s][y][lx] * sP[s][y][lx + 1];
                    acc += sWx[s][y][lx - 1] * sP[s][y][lx - 1];
                    acc += sWy[s][y][lx] * sP[s][y + 1][lx];
                    acc += sWy[s][y - 1][lx] * sP[s][y - 1][lx];
                    acc += rwz[q][1 + h] * p[q][h];
                    acc += rwz[q][2 + h] * p[q][h + 2];
                    const T pc = p[q][1 + h];
                    const T pn = pc + omega * (acc * rinv[q][1 + h] - pc);
                    p[q][1 + h] = pn;
                    sP[s][y][lx] = pn;
                }
            }
        }
        const int mo = k - R;
        if (mo >= k0 && mo < k1) {
#pragma unroll
            for (int q = 0; q < RY; ++q)
                if (d[q] >= R && in_dom[q]) Pout[col[q] + plane * mo] = p[q][R];
        }
        s0 = (s0 + 1 == NSLOT) ? 0 : s0 + 1;
    }
}

// ---- cross-lane x neighbours -------------------------------------------------------------------
// A tile row is one 64-lane wavefront (lane = x offset), so the cells left
// and right of lane i sit in lanes i-1 and i+1: DPP wave shift right / left
// by one lane moves a register between them (lanes 0 and 63 receive 0 through
// bound_ctrl; the tile's outer columns are halo and never updated).
__device__ __forceinline__ int lane_shr1(int v) {  // lane i <- lane i-1
    return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int lane_shl1(int v) {  // lane i <- lane i+1
    return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(lane_shr1(__float_as_int(v)));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(lane_shl1(__float_as_int(v)));
}
__device__ __forceinline__ double from_left(double v) {
    const long long b = __double_as_longlong(v);
    const long long lo = (unsigned)lane_shr1((int)(b & 0xffffffffll));
    const long long hi = lane_shr1((int)(b >> 32));
    return __longlong_as_double((hi << 32) | lo);
}
__device__ __forceinline__ double from_right(double v) {
    const long long b = __double_as_longlong(v);
    const long long lo = (unsigned)lane_shl1((int)(b & 0xffffffffll));
    const long long hi = lane_shl1((int)(b >> 32));
    return __longlong_as_double((hi << 32) | lo);
}

// Variant of rb_zmarch_tile with the x neighbours exchanged across lanes:
// only P and Wy go through LDS (two plane rings instead of three), each
// thread keeps its own Wx and Wy of the in-flight planes in registers, and
// P(x+-1) and Wx(x-1) come from the adjacent lanes.  Same operands in the
// same order as rb_zmarch_tile; the smaller LDS ring allows taller tiles at
// the same occupancy.
template <typename T, int NS, int BYT, int RY>
__device__ __forceinline__ void
rb_zmarch_tile_xl(const ZGeo &g, int tile, const T *__restrict__ Pin,
                  T *__restrict__ Pout, const T *__restrict__ Wx,
                  const T *__restrict__ Wy, const T *__restrict__ Wz,
                  const T *__restrict__ invD, const T *__restrict__ rhs, T omega,
                  int zero_init) {
    constexpr int R = 2 * NS;
    constexpr int BX = 64;
    constexpr int BY = BYT * RY;
    constexpr int TX = BX - 2 * R;
    constexpr int TY = BY - 2 * R;
    constexpr int NSLOT = R + 2;
    static_assert(TY >= 1, "tile too small");
    __shared__ T sP[NSLOT][BY][BX];
    __shared__ T sWy[NSLOT][BY][BX];

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);
    const int ty = zc - zc2 * g.tiles_y;
    const int lx = threadIdx.x;
    const int i = tx * TX - R + lx;
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const long plane = (long)g.nx * g.ny;
    const int dxe = min(lx, BX - 1 - lx);

    int ly[RY], d[RY], par[RY];
    bool in_dom[RY];
    long col[RY];
    T p[RY][R + 2], rwz[RY][R + 2], rinv[RY][R + 1], rb[RY][R + 1];
    T rwx[RY][R + 1], rwy[RY][R + 1];
    T np[RY], ni[RY], nb[RY], nwz[RY], nwx[RY], nwy[RY];
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        ly[q] = threadIdx.y + q * BYT;
        const int j = ty * TY - R + ly[q];
        in_dom[q] = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
        d[q] = min(dxe, min(ly[q], BY - 1 - ly[q]));
        col[q] = in_dom[q] ? (long)i + (long)g.nx * j : 0;
        par[q] = (i + j) & 1;
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
            p[q][r] = T(0);
            rwz[q][r] = T(0);
        }
#pragma unroll
        for (int r = 0; r < R + 1; ++r) {
            rinv[q][r] = rb[q][r] = rwx[q][r] = rwy[q][r] = T(0);
        }
        np[q] = ni[q] = nb[q] = nwz[q] = nwx[q] = nwy[q] = T(0);
        if (in_dom[q] && kl0 <= kl1) {
            const long o = col[q] + plane * kl0;
            np[q] = zero_init ? T(0) : Pin[o];
            ni[q] = invD[o];
            nb[q] = rhs[o];
            nwz[q] = Wz[o];
            nwx[q] = Wx[o];
            nwy[q] = Wy[o];
        }
    }

    int s0 = kl0 % NSLOT;
    for (int k = kl0; k <= k1 + R - 1; ++k) {
#pragma unroll
        for (int q = 0; q < RY; ++q) {
            const T p0 = np[q], i0 = ni[q], b0 = nb[q], wz0 = nwz[q];
            const T wx0 = nwx[q], wy0 = nwy[q];
            np[q] = ni[q] = nb[q] = nwz[q] = nwx[q] = nwy[q] = T(0);
            if (in_dom[q] && k + 1 <= kl1) {
                const long o = col[q] + plane * (k + 1);
                np[q] = zero_init ? T(0) : Pin[o];
                ni[q] = invD[o];
                nb[q] = rhs[o];
                nwz[q] = Wz[o];
                nwx[q] = Wx[o];
                nwy[q] = Wy[o];
            }
#pragma unroll
            for (int r = R + 1; r > 0; --r) {
                p[q][r] = p[q][r - 1];
                rwz[q][r] = rwz[q][r - 1];
            }
#pragma unroll
            for (int r = R; r > 0; --r) {
                rinv[q][r] = rinv[q][r - 1];
                rb[q][r] = rb[q][r - 1];
                rwx[q][r] = rwx[q][r - 1];
                rwy[q][r] = rwy[q][r - 1];
            }
            p[q][0] = p0;
            rwz[q][0] = wz0;
            rinv[q][0] = i0;
            rb[q][0] = b0;
            rwx[q][0] = wx0;
            rwy[q][0] = wy0;
            sP[s0][ly[q]][lx] = p0;
            sWy[s0][ly[q]][lx] = wy0;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < R; ++h) {
            const int m = k - 1 - h;
            int s = s0 - 1 - h;
            if (s < 0) s += NSLOT;
            const bool mok = m >= kl0 && m <= kl1;
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                // all lanes take part in the exchange; x neighbours have the
                // other colour, so phase h does not change what is read
                const T pl = from_left(p[q][1 + h]);
                const T pr = from_right(p[q][1 + h]);
                const T wxl = from_left(rwx[q][1 + h]);
                if (mok && d[q] >= h + 1 && in_dom[q] &&
                    ((par[q] + m) & 1) == (h & 1) && rinv[q][1 + h] > T(0)) {
                    const int y = ly[q];
                    T acc = rb[q][1 + h];
                    acc += rwx[q][1 + h] * pr;
                    acc += wxl * pl;
                    acc += rwy[q][1 + h] * sP[s][y + 1][lx];
                    acc += sWy[s][y - 1][lx] * sP[s][y - 1][lx];
                    acc += rwz[q][1 + h] * p[q][h];
                    acc += rwz[q][2 + h] * p[q][h + 2];
                    const T pc = p[q][1 + h];
                    const T pn = pc + omega * (acc * rinv[q][1 + h] - pc);
                    p[q][1 + h] = pn;
                    sP[s][y][lx] = pn;
                }
            }
        }
        const int mo = k - R;
        if (mo >= k0 && mo < k1) {
#pragma unroll
            for (int q = 0; q < RY; ++q)
                if (d[q] >= R && in_dom[q]) Pout[col[q] + plane * mo] = p[q][R];
        }
        s0 = (s0 + 1 == NSLOT) ? 0 : s0 + 1;
    }
}

// Raw buffer loads (gfx9 descriptor: base, stride 0, size in bytes, word 3
// 0x00020000).  An offset at or past the size returns 0 without a memory
// access, which zero-fills out-of-domain lanes and planes without branches.
constexpr uint32_t kBufMax = 1u << 30;   // field size limit of the buffer forms
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const T *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
// A store at an offset beyond the field is discarded.
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0, 0);
}
typedef unsigned int tse_u32x2 __attribute__((__vector_size__(8)));
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tse_u32x2, v), r, (int)off, 0, 0);
}

// Variant of rb_zmarch_tile_xl with the z-loop unrolled by the ring length
// L = 2*NS + 2: plane m sits in slot (m - kl0) mod L of every per-thread ring
// and of the LDS rings, so the rings rotate by renaming instead of register
// moves.  The coefficient planes are loaded straight into their slots one
// step ahead (the slot's previous plane, k+1-L, is no longer read), P and Wz
// through one prefetch register each.  Loads are raw buffer loads at one
// shared 32-bit byte offset per step; lanes outside the grid and planes
// past the tile's last plane use offsets beyond the field and read 0 (fields
// up to kBufMax bytes: in-grid offsets < 2^30, outside-lane column offset
// 2^30, past-the-end plane offset 2^31, so no sum wraps); planes outside
// [kl0, kl1] therefore hold invD = 0 and are skipped by the Dirichlet test,
// which makes a separate plane-range test unnecessary.  The output store
// goes through a buffer store whose offset is beyond the field for lanes and
// planes that must not be written.  Phase h at step k updates plane k-1-h in
// colour h&1, i.e. the cells with (i+j+k-1) even whatever h is, so the
// colour test is a lane constant that alternates between steps.  Same
// operands in the same order as rb_zmarch_tile_xl: bit-identical results.
// LDS rings of the variant-2/3 tiles: [ring slot][row + 1][x]; rows 0 and
// BY+1 are zero guards for the y neighbours of boundary rows in tiles whose
// edge is the grid boundary.
template <typename T, int NS, int BY>
using ZRing = T[2 * NS + 2][BY + 2][64];

template <typename T, int NS, int BYT, int RY>
__device__ __forceinline__ void
rb_zmarch_tile_xr_in(const ZGeo &g, int tile, const T *__restrict__ Pin,
                     T *__restrict__ Pout, const T *__restrict__ Wx,
                     const T *__restrict__ Wy, const T *__restrict__ Wz,
                     const T *__restrict__ invD, const T *__restrict__ rhs, T omega,
                     int zero_init, ZRing<T, NS, BYT * RY> &sP,
                     ZRing<T, NS, BYT * RY> &sWy) {
    constexpr int R = 2 * NS;
    constexpr int BX = 64;
    constexpr int BY = BYT * RY;
    constexpr int TX = BX - 2 * R;
    constexpr int TY = BY - 2 * R;
    constexpr int L = R + 2;
    static_assert(TY >= 1, "tile too small");

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);
    const int ty = zc - zc2 * g.tiles_y;
    const int lx = threadIdx.x;
    const Span sx = tile_span(g.edge, tx, g.tiles_x, g.nx, BX, R);
    const Span sy = tile_span(g.edge, ty, g.tiles_y, g.ny, BY, R);
    const int i = sx.x0 + lx;
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const int klast = k1 + R - 1;
    const uint32_t pbytes = (uint32_t)g.nx * (uint32_t)g.ny * (uint32_t)sizeof(T);
    const uint32_t fbytes = pbytes * (uint32_t)g.nz;
    const auto rP = buf_rsrc(Pin, zero_init ? 0u : fbytes);
    const auto rWx = buf_rsrc(Wx, fbytes), rWy = buf_rsrc(Wy, fbytes);
    const auto rWz = buf_rsrc(Wz, fbytes), rI = buf_rsrc(invD, fbytes);
    const auto rB = buf_rsrc(rhs, fbytes);
    const auto rOut = buf_rsrc(Pout, fbytes);
    const int dxe = min(sx.lfree ? kFar : lx, sx.hfree ? kFar : BX - 1 - lx);

    int ly[RY];
    bool in_dom[RY], cm[RY], dm[RY][R];
    uint32_t ocol[RY], oout[RY];
    T p[RY][L], rwz[RY][L], rinv[RY][L], rb[RY][L], rwx[RY][L], rwy[RY][L];
    T np[RY], nwz[RY];
    const uint32_t ok0 = kl0 <= kl1 ? (uint32_t)kl0 * pbytes : 0x80000000u;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        ly[q] = threadIdx.y + q * BYT;
        const int j = sy.x0 + ly[q];
        in_dom[q] = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
        const int d = min(dxe, min(sy.lfree ? kFar : ly[q],
                                   sy.hfree ? kFar : BY - 1 - ly[q]));
#pragma unroll
        for (int h = 0; h < R; ++h) dm[q][h] = in_dom[q] && d >= h + 1;
        // colour test of the steps with (k - kl0) even
        cm[q] = ((i + j + kl0 - 1) & 1) == 0;
        ocol[q] = in_dom[q] ? (uint32_t)(i + g.nx * j) * (uint32_t)sizeof(T) : kBufMax;
        // the cells this tile owns (valid after all R phases)
        const bool own = in_dom[q] && i >= sx.lo && i < sx.hi && j >= sy.lo && j < sy.hi;
        oout[q] = own ? ocol[q] : kBufMax;
#pragma unroll
        for (int r = 1; r < L; ++r)
            p[q][r] = rwz[q][r] = rinv[q][r] = rb[q][r] = rwx[q][r] = rwy[q][r] = T(0);
        const uint32_t o = ocol[q] + ok0;
        np[q] = buf_ld(rP, o, T());
        nwz[q] = buf_ld(rWz, o, T());
        rinv[q][0] = buf_ld(rI, o, T());
        rb[q][0] = buf_ld(rB, o, T());
        rwx[q][0] = buf_ld(rWx, o, T());
        rwy[q][0] = buf_ld(rWy, o, T());
        p[q][0] = rwz[q][0] = T(0);
    }

    if (threadIdx.y == 0)
#pragma unroll
        for (int u = 0; u < L; ++u) sP[u][0][lx] = sWy[u][0][lx] = T(0);
    if (threadIdx.y == BYT - 1)
#pragma unroll
        for (int u = 0; u < L; ++u) sP[u][BY + 1][lx] = sWy[u][BY + 1][lx] = T(0);
    for (int kb = kl0; kb <= klast; kb += L) {
#pragma unroll
        for (int u = 0; u < L; ++u) {
            const int k = kb + u;
            if (k > klast) break;
            const int un = (u + 1) % L;
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                p[q][u] = np[q];
                rwz[q][u] = nwz[q];
                sP[u][ly[q] + 1][lx] = np[q];
                sWy[u][ly[q] + 1][lx] = rwy[q][u];
            }
            // plane k+1 (reads 0 past kl1 and outside the grid)
            const uint32_t ok = k + 1 <= kl1 ? (uint32_t)(k + 1) * pbytes : 0x80000000u;
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                const uint32_t o = ocol[q] + ok;
                np[q] = buf_ld(rP, o, T());
                nwz[q] = buf_ld(rWz, o, T());
                rinv[q][un] = buf_ld(rI, o, T());
                rb[q][un] = buf_ld(rB, o, T());
                rwx[q][un] = buf_ld(rWx, o, T());
                rwy[q][un] = buf_ld(rWy, o, T());
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < R; ++h) {
                const int sm = (u - 1 - h + 2 * L) % L;     // plane m = k-1-h
                const int sa = (u - h + 2 * L) % L;         // plane m + 1
                const int sb = (u - 2 - h + 2 * L) % L;     // plane m - 1
#pragma unroll
                for (int q = 0; q < RY; ++q) {
                    const T pl = from_left(p[q][sm]);
                    const T pr = from_right(p[q][sm]);
                    const T wxl = from_left(rwx[q][sm]);
                    const bool col_ok = (u & 1) ? !cm[q] : cm[q];
                    if (dm[q][h] && col_ok && rinv[q][sm] > T(0)) {
                        const int y = ly[q] + 1;   // LDS row
                        T acc = rb[q][sm];
                        acc += rwx[q][sm] * pr;
                        acc += wxl * pl;
                        acc += rwy[q][sm] * sP[sm][y + 1][lx];
                        acc += sWy[sm][y - 1][lx] * sP[sm][y - 1][lx];
                        acc += rwz[q][sm] * p[q][sa];
                        acc += rwz[q][sb] * p[q][sb];
                        const T pc = p[q][sm];
                        const T pn = pc + omega * (acc * rinv[q][sm] - pc);
                        p[q][sm] = pn;
                        sP[sm][y][lx] = pn;
                    }
                }
            }
            const int mo = k - R;
            const uint32_t oo = mo >= k0 && mo < k1 ? (uint32_t)mo * pbytes : 0x80000000u;
#pragma unroll
            for (int q = 0; q < RY; ++q)
                buf_st(p[q][(u - R + 2 * L) % L], rOut, oout[q] + oo);
        }
    }
}

template <typename T, int NS, int BYT, int RY>
__device__ __forceinline__ void
rb_zmarch_tile_xr(const ZGeo &g, int tile, const T *__restrict__ Pin,
                  T *__restrict__ Pout, const T *__restrict__ Wx,
                  const T *__restrict__ Wy, const T *__restrict__ Wz,
                  const T *__restrict__ invD, const T *__restrict__ rhs, T omega,
                  int zero_init) {
    __shared__ ZRing<T, NS, BYT * RY> sP, sWy;
    rb_zmarch_tile_xr_in<T, NS, BYT, RY>(g, tile, Pin, Pout, Wx, Wy, Wz, invD, rhs,
                                         omega, zero_init, sP, sWy);
}

// Variant 3 (FP32, even nx): each thread holds an x-pair of cells (x, x+1,
// x even) of one row, and each wavefront two rows of equal parity (block
// row y of thread (lane, ty) is 4*(ty/2) + (ty&1) + 2*(lane/32)).  Phase h
// at step k updates the cells with (i+j+k-1) even, so in a plane step every
// phase updates the same element of every pair: element (j+k-1)&1.  The z
// loop is unrolled by the even ring length, so that element is a compile-
// time function of the row parity of the wavefront (template Q, chosen by
// a uniform branch) and of the unrolled step; every lane updates one cell in
// every phase, where the one-cell-per-lane variants leave half the lanes
// idle.  Pairs are read and written with 8-byte buffer accesses.  The x
// neighbour inside the pair is the other element, the one across the pair
// boundary comes from the adjacent lane (DPP); lanes 0 and 32 start a row,
// so the shifted-in west face coefficient is zeroed there.  Same operands
// in the same order as the other variants: bit-identical results.
typedef float tse_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ tse_f2 buf_ld2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(tse_f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void buf_st2(tse_f2 v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tse_u32x2, v), r, (int)off, 0, 0);
}

template <int NS, int BYT, int Q, bool BF>
__device__ __forceinline__ void
rb_zmarch_x2_march(int ly, int ip, int k0, int k1, int kl0, int kl1, uint32_t pbytes,
                   uint32_t ocol, uint32_t oout, const bool (&dm)[2][2 * NS],
                   __amdgpu_buffer_rsrc_t rP, __amdgpu_buffer_rsrc_t rWx,
                   __amdgpu_buffer_rsrc_t rWy, __amdgpu_buffer_rsrc_t rWz,
                   __amdgpu_buffer_rsrc_t rI, __amdgpu_buffer_rsrc_t rB,
                   __amdgpu_buffer_rsrc_t rOut, float omega,
                   float (&sP)[2 * NS + 2][2 * BYT + 2][64],
                   float (&sWy)[2 * NS + 2][2 * BYT + 2][64]) {
    constexpr int R = 2 * NS;
    constexpr int L = R + 2;
    const int klast = k1 + R - 1;
    const int y = ly + 1;             // LDS row
    // LDS row r holds two 32-word blocks, element e of pair ip at word
    // 32 * (e ^ ((r >> 1) & 1)) + ip: a half-wavefront touches 32 consecutive
    // words and the two halves (rows r and r + 2) use different blocks
    const int cm1[2] = {(((0 ^ ((y - 1) >> 1)) & 1) << 5) + ip, (((1 ^ ((y - 1) >> 1)) & 1) << 5) + ip};
    const int c0[2] = {(((0 ^ (y >> 1)) & 1) << 5) + ip, (((1 ^ (y >> 1)) & 1) << 5) + ip};
    const int cp1[2] = {(((0 ^ ((y + 1) >> 1)) & 1) << 5) + ip, (((1 ^ ((y + 1) >> 1)) & 1) << 5) + ip};
    float p[2][L], rwz[2][L], rinv[2][L], rb[2][L], rwx[2][L], rwy[2][L];
    tse_f2 np, nwz;
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int r = 0; r < L; ++r)
            p[e][r] = rwz[e][r] = rinv[e][r] = rb[e][r] = rwx[e][r] = rwy[e][r] = 0.f;
    {
        const uint32_t o = ocol + (kl0 <= kl1 ? (uint32_t)kl0 * pbytes : 0x80000000u);
        np = buf_ld2(rP, o);
        nwz = buf_ld2(rWz, o);
        const tse_f2 a = buf_ld2(rI, o), b = buf_ld2(rB, o), c = buf_ld2(rWx, o),
                     d = buf_ld2(rWy, o);
        rinv[0][0] = a.x; rinv[1][0] = a.y;
        rb[0][0] = b.x;   rb[1][0] = b.y;
        rwx[0][0] = c.x;  rwx[1][0] = c.y;
        rwy[0][0] = d.x;  rwy[1][0] = d.y;
    }
    const bool row_start = ip == 0;   // lanes 0 and 32
    for (int kb = kl0; kb <= klast; kb += L) {
#pragma unroll
        for (int u = 0; u < L; ++u) {
            const int k = kb + u;
            if (k > klast) break;
            const int un = (u + 1) % L;
            p[0][u] = np.x;
            p[1][u] = np.y;
            rwz[0][u] = nwz.x;
            rwz[1][u] = nwz.y;
            sP[u][y][c0[0]] = np.x;
            sP[u][y][c0[1]] = np.y;
            sWy[u][y][c0[0]] = rwy[0][u];
            sWy[u][y][c0[1]] = rwy[1][u];
            {   // plane k+1 (reads 0 past kl1 and outside the grid)
                const uint32_t o =
                    ocol + (k + 1 <= kl1 ? (uint32_t)(k + 1) * pbytes : 0x80000000u);
                np = buf_ld2(rP, o);
                nwz = buf_ld2(rWz, o);
                const tse_f2 a = buf_ld2(rI, o), b = buf_ld2(rB, o), c = buf_ld2(rWx, o),
                             d = buf_ld2(rWy, o);
                rinv[0][un] = a.x; rinv[1][un] = a.y;
                rb[0][un] = b.x;   rb[1][un] = b.y;
                rwx[0][un] = c.x;  rwx[1][un] = c.y;
                rwy[0][un] = d.x;  rwy[1][un] = d.y;
            }
            __syncthreads();
            const int eact = (Q + u) & 1;   // compile-time after unrolling
#pragma unroll
            for (int h = 0; h < R; ++h) {
                const int sm = (u - 1 - h + 2 * L) % L;     // plane m = k-1-h
                const int sa = (u - h + 2 * L) % L;         // plane m + 1
                const int sb = (u - 2 - h + 2 * L) % L;     // plane m - 1
                float pl, pr, wxl;
                if (eact == 0) {
                    pl = from_left(p[1][sm]);
                    pr = p[1][sm];
                    const float w = from_left(rwx[1][sm]);
                    wxl = row_start ? 0.f : w;
                } else {
                    pl = p[0][sm];
                    pr = from_right(p[0][sm]);
                    wxl = rwx[0][sm];
                }
                const bool go = dm[eact][h] && rinv[eact][sm] > 0.f;
                auto update = [&]() {
                    float acc = rb[eact][sm];
                    acc += rwx[eact][sm] * pr;
                    acc += wxl * pl;
                    acc += rwy[eact][sm] * sP[sm][y + 1][cp1[eact]];
                    acc += sWy[sm][y - 1][cm1[eact]] * sP[sm][y - 1][cm1[eact]];
                    acc += rwz[eact][sm] * p[eact][sa];
                    acc += rwz[eact][sb] * p[eact][sb];
                    const float pc = p[eact][sm];
                    return pc + omega * (acc * rinv[eact][sm] - pc);
                };
                if constexpr (BF) {
                    // branch-free: every lane computes and cells that must not
                    // change keep their value, so a phase's LDS reads need not
                    // wait for the branch of the phase before (more registers)
                    const float pn = update();
                    const float pv = go ? pn : p[eact][sm];
                    p[eact][sm] = pv;
                    sP[sm][y][c0[eact]] = pv;
                } else if (go) {
                    const float pn = update();
                    p[eact][sm] = pn;
                    sP[sm][y][c0[eact]] = pn;
                }
            }
            const int mo = k - R;
            const uint32_t oo = mo >= k0 && mo < k1 ? (uint32_t)mo * pbytes : 0x80000000u;
            const int so = (u - R + 2 * L) % L;
            tse_f2 out;
            out.x = p[0][so];
            out.y = p[1][so];
            buf_st2(out, rOut, oout + oo);
        }
    }
}

template <int NS, int BYT, bool BF>
__device__ __forceinline__ void
rb_zmarch_tile_x2_in(const ZGeo &g, int tile, const float *__restrict__ Pin,
                     float *__restrict__ Pout, const float *__restrict__ Wx,
                     const float *__restrict__ Wy, const float *__restrict__ Wz,
                     const float *__restrict__ invD, const float *__restrict__ rhs,
                     float omega, int zero_init, ZRing<float, NS, 2 * BYT> &sP,
                     ZRing<float, NS, 2 * BYT> &sWy) {
    constexpr int R = 2 * NS;
    constexpr int BX = 64;
    constexpr int BY = 2 * BYT;
    constexpr int L = R + 2;
    static_assert(BY - 2 * R >= 1, "tile too small");

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);
    const int ty = zc - zc2 * g.tiles_y;
    const Span sx = tile_span(g.edge, tx, g.tiles_x, g.nx, BX, R);
    const Span sy = tile_span(g.edge, ty, g.tiles_y, g.ny, BY, R);
    const int ip = threadIdx.x & 31;
    const int ly = ((threadIdx.y >> 1) << 2) + (threadIdx.y & 1) + ((threadIdx.x >> 5) << 1);
    const int i = sx.x0 + 2 * ip;     // even (nx even, so every tile starts at an even x)
    const int j = sy.x0 + ly;
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const uint32_t pbytes = (uint32_t)g.nx * (uint32_t)g.ny * 4u;
    const uint32_t fbytes = pbytes * (uint32_t)g.nz;
    const auto rP = buf_rsrc(Pin, zero_init ? 0u : fbytes);
    const auto rWx = buf_rsrc(Wx, fbytes), rWy = buf_rsrc(Wy, fbytes);
    const auto rWz = buf_rsrc(Wz, fbytes), rI = buf_rsrc(invD, fbytes);
    const auto rB = buf_rsrc(rhs, fbytes);
    const auto rOut = buf_rsrc(Pout, fbytes);
    const bool in_dom = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
    const int dy = min(sy.lfree ? kFar : ly, sy.hfree ? kFar : BY - 1 - ly);
    bool dm[2][R];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int xt = 2 * ip + e;
        const int d = min(dy, min(sx.lfree ? kFar : xt, sx.hfree ? kFar : BX - 1 - xt));
#pragma unroll
        for (int h = 0; h < R; ++h) dm[e][h] = in_dom && d >= h + 1;
    }
    const uint32_t ocol = in_dom ? (uint32_t)(i + g.nx * j) * 4u : kBufMax;
    // pairs owned by this tile (lo and hi are even)
    const bool own = in_dom && i >= sx.lo && i < sx.hi && j >= sy.lo && j < sy.hi;
    const uint32_t oout = own ? ocol : kBufMax;
    if (threadIdx.y == 0)
#pragma unroll
        for (int u = 0; u < L; ++u) sP[u][0][threadIdx.x] = sWy[u][0][threadIdx.x] = 0.f;
    if (threadIdx.y == BYT - 1)
#pragma unroll
        for (int u = 0; u < L; ++u)
            sP[u][BY + 1][threadIdx.x] = sWy[u][BY + 1][threadIdx.x] = 0.f;
    // active element of the first step: (j + kl0 - 1) & 1, the same for the
    // whole wavefront (its rows have one parity)
    const int q = __builtin_amdgcn_readfirstlane((j + kl0 - 1) & 1);
    if (q == 0)
        rb_zmarch_x2_march<NS, BYT, 0, BF>(ly, ip, k0, k1, kl0, kl1, pbytes, ocol, oout,
                                           dm, rP, rWx, rWy, rWz, rI, rB, rOut, omega,
                                           sP, sWy);
    else
        rb_zmarch_x2_march<NS, BYT, 1, BF>(ly, ip, k0, k1, kl0, kl1, pbytes, ocol, oout,
                                           dm, rP, rWx, rWy, rWz, rI, rB, rOut, omega,
                                           sP, sWy);
}

// BF: branch-free phases (the single-system kernel: 485 against 503 us per
// 512x512x256 launch, with 8 bytes of scratch); the batch kernels keep the
// branch, where the branch-free form spills more and runs slower.
template <int NS, int BYT, bool BF>
__device__ __forceinline__ void
rb_zmarch_tile_x2(const ZGeo &g, int tile, const float *__restrict__ Pin,
                  float *__restrict__ Pout, const float *__restrict__ Wx,
                  const float *__restrict__ Wy, const float *__restrict__ Wz,
                  const float *__restrict__ invD, const float *__restrict__ rhs,
                  float omega, int zero_init) {
    __shared__ ZRing<float, NS, 2 * BYT> sP, sWy;
    rb_zmarch_tile_x2_in<NS, BYT, BF>(g, tile, Pin, Pout, Wx, Wy, Wz, invD, rhs, omega,
                                      zero_init, sP, sWy);
}

// Variant 5 (FP32 arithmetic, even nx): the cell-pair z-march of variant 3
// on an inner operator stored in 8 bytes per cell.  Each pair (x, x+1) has a
// 16-byte record of four 32-bit words: Wx, Wy, Wz and S = V R (the storage
// term) of the pair, each as two bfloat16 values (element 0 in the low
// half), read with one 16-byte buffer load; the rhs and the unknown stay
// FP32.  The diagonal is summed in the kernel from the same rounded values,
// D' = S' + sum of the cell's six rounded faces, and S' = +inf marks a
// Dirichlet cell (1/D' = 0).  The inner operator
//   A' = sum_f W'_f L_f + diag(S'),  W'_f = W_f (1 + d_f),  S' = S (1 + d_S),
// |d| <= u = 2^-8 (bfloat16's unit roundoff), is a graph Laplacian with
// perturbed face weights plus a perturbed positive diagonal, so
// (1 - u) A <= A' <= (1 + u) A: a correction cycle that solved A' e = r
// exactly would leave at most u / (1 - u) of the error (energy norm), and a
// cycle of SOR sweeps leaves the SOR contraction plus that.  The defect
// correction (FP64 residual with the FP64 operator, same stop test)
// therefore reaches the same tolerance with the same sweep count while
// cycles reduce the residual by less than ~2 decades (adaptive cycles aim at
// most that far).  A launch
// moves 8 + 4 + 4 + 4 = 20 bytes per cell (28 for the FP32 operator).
//
// No validity mask: in the correction equation every cell whose 1/diagonal
// is 0 (Dirichlet cells, lanes outside the grid, planes outside the chunk)
// holds e = 0, and the update leaves it at (1 - omega) * 0 = 0; halo cells
// near an interior tile edge are updated with stale neighbours, which cannot
// reach the owned cells (a cell at distance d from an interior tile edge is
// correct after phase h iff d >= h + 1, whatever the cells nearer the edge
// hold, as long as they stay finite: the LDS rings start at zero).  So every
// lane updates in every phase with no select.
__device__ __forceinline__ uint16_t bf16_rne(double w) {  // w >= 0, finite
    uint32_t u = __float_as_uint((float)w);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ double bf16_value(uint16_t h) {
    return (double)__uint_as_float((uint32_t)h << 16);
}
constexpr uint16_t kBf16Inf = 0x7f80u;
template <int E>
__device__ __forceinline__ float bf_el(uint32_t v) {  // element E of a packed pair
    return E ? __uint_as_float(v & 0xffff0000u) : __uint_as_float(v << 16);
}
template <int E>
__device__ __forceinline__ float h_el(uint32_t v) {  // element E of a packed FP16 pair
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(E ? v >> 16 : v & 0xffffu));
}
typedef unsigned int tse_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ tse_u32x4 buf_ld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
}

// LDS ring of packed Wy pairs: row r of a slot sits in 64-word group r / 2,
// half (r & 1) ^ ((r / 2) & 1), so rows r and r + 2 (the two halves of a
// wavefront) fall into different 32-word halves: no bank conflicts.
template <int NS, int BYT>
using WyRing = uint32_t[2 * NS + 2][BYT + 1][64];
__device__ __forceinline__ int wy_word(int r, int ip) {
    return (((r & 1) ^ ((r >> 1) & 1)) << 5) + ip;
}

// RB16 (variant 7): the rhs is read as one packed FP16 pair per thread (4
// bytes at ohcol + k * hbytes, scaled by a power of two chosen by the
// residual pass) and kept packed in its register ring; otherwise as an FP32
// pair at the unknown's offsets.
template <int NS, int BYT, int Q, bool RB16>
__device__ __forceinline__ void
rb_zmarch_x5_march(int ly, int ip, int k0, int k1, int kl0, int kl1, uint32_t pbytes,
                   uint32_t rbytes, uint32_t hbytes, uint32_t ocol, uint32_t orcol,
                   uint32_t ohcol, uint32_t oout,
                   __amdgpu_buffer_rsrc_t rP, __amdgpu_buffer_rsrc_t rRec,
                   __amdgpu_buffer_rsrc_t rB, __amdgpu_buffer_rsrc_t rOut, float omega,
                   float (&sP)[2 * NS + 2][2 * BYT + 2][64], WyRing<NS, BYT> &sWy) {
    constexpr int R = 2 * NS;
    constexpr int L = R + 2;
    const int klast = k1 + R - 1;
    const int y = ly + 1;             // LDS row
    const int cm1[2] = {(((0 ^ ((y - 1) >> 1)) & 1) << 5) + ip, (((1 ^ ((y - 1) >> 1)) & 1) << 5) + ip};
    const int c0[2] = {(((0 ^ (y >> 1)) & 1) << 5) + ip, (((1 ^ (y >> 1)) & 1) << 5) + ip};
    const int cp1[2] = {(((0 ^ ((y + 1) >> 1)) & 1) << 5) + ip, (((1 ^ ((y + 1) >> 1)) & 1) << 5) + ip};
    const int gw0 = y >> 1, ww0 = wy_word(y, ip);            // own row in the Wy ring
    const int gwm = (y - 1) >> 1, wwm = wy_word(y - 1, ip);  // row below
    float p[2][L], rinv[2][L], rb[2][RB16 ? 1 : L];
    uint32_t rwx[L], rwy[L], rwz[L], rbh[RB16 ? L : 1];
    tse_f2 np;
    uint32_t nwz, nsd;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        rwx[r] = rwy[r] = rwz[r] = 0u;
#pragma unroll
        for (int e = 0; e < 2; ++e) p[e][r] = rinv[e][r] = 0.f;
        if constexpr (RB16) rbh[r] = 0u;
        else rb[0][r] = rb[1][r] = 0.f;
    }
    {
        const bool in = kl0 <= kl1;
        const uint32_t o = ocol + (in ? (uint32_t)kl0 * pbytes : 0x80000000u);
        const uint32_t orr = orcol + (in ? (uint32_t)kl0 * rbytes : 0x80000000u);
        np = buf_ld2(rP, o);
        const tse_u32x4 w = buf_ld4(rRec, orr);
        rwx[0] = w.x;
        rwy[0] = w.y;
        nwz = w.z;
        nsd = w.w;
        if constexpr (RB16) {
            const uint32_t oh = ohcol + (in ? (uint32_t)kl0 * hbytes : 0x80000000u);
            rbh[0] = __builtin_amdgcn_raw_buffer_load_b32(rB, (int)oh, 0, 0);
        } else {
            const tse_f2 b = buf_ld2(rB, o);
            rb[0][0] = b.x;
            rb[1][0] = b.y;
        }
    }
    const bool row_start = ip == 0;   // lanes 0 and 32
    for (int kb = kl0; kb <= klast; kb += L) {
#pragma unroll
        for (int u = 0; u < L; ++u) {
            const int k = kb + u;
            if (k > klast) break;
            const int un = (u + 1) % L;
            p[0][u] = np.x;
            p[1][u] = np.y;
            rwz[u] = nwz;
            const uint32_t sd = nsd;   // storage term of plane k
            sP[u][y][c0[0]] = np.x;
            sP[u][y][c0[1]] = np.y;
            sWy[u][gw0][ww0] = rwy[u];
            {   // plane k+1 (reads 0 past kl1 and outside the grid)
                const bool in = k + 1 <= kl1;
                const uint32_t o = ocol + (in ? (uint32_t)(k + 1) * pbytes : 0x80000000u);
                const uint32_t orr =
                    orcol + (in ? (uint32_t)(k + 1) * rbytes : 0x80000000u);
                np = buf_ld2(rP, o);
                const tse_u32x4 w = buf_ld4(rRec, orr);
                rwx[un] = w.x;
                rwy[un] = w.y;
                nwz = w.z;
                nsd = w.w;
                if constexpr (RB16) {
                    const uint32_t oh =
                        ohcol + (in ? (uint32_t)(k + 1) * hbytes : 0x80000000u);
                    rbh[un] = __builtin_amdgcn_raw_buffer_load_b32(rB, (int)oh, 0, 0);
                } else {
                    const tse_f2 b = buf_ld2(rB, o);
                    rb[0][un] = b.x;
                    rb[1][un] = b.y;
                }
            }
            __syncthreads();
            {   // 1/diagonal of plane k (first used by phase 0 at step k+1):
                // S' + the six rounded faces; S' = +inf (Dirichlet) gives 0,
                // and so does a zero plane (outside the chunk or the grid)
                const uint32_t wyb = sWy[u][gwm][wwm];
                const uint32_t wxw = (uint32_t)lane_shr1((int)rwx[u]);
                const int ub = (u + L - 1) % L;
                const float west0 = row_start ? 0.f : bf_el<1>(wxw);
                const float d0 = ((bf_el<0>(sd) + bf_el<0>(rwx[u])) + (west0 + bf_el<0>(rwy[u]))) +
                                 ((bf_el<0>(wyb) + bf_el<0>(rwz[u])) + bf_el<0>(rwz[ub]));
                const float d1 = ((bf_el<1>(sd) + bf_el<1>(rwx[u])) +
                                  (bf_el<0>(rwx[u]) + bf_el<1>(rwy[u]))) +
                                 ((bf_el<1>(wyb) + bf_el<1>(rwz[u])) + bf_el<1>(rwz[ub]));
                rinv[0][u] = d0 > 0.f ? __builtin_amdgcn_rcpf(d0) : 0.f;
                rinv[1][u] = d1 > 0.f ? __builtin_amdgcn_rcpf(d1) : 0.f;
            }
            const int eact = (Q + u) & 1;   // compile-time after unrolling
#pragma unroll
            for (int h = 0; h < R; ++h) {
                const int sm = (u - 1 - h + 2 * L) % L;     // plane m = k-1-h
                const int sa = (u - h + 2 * L) % L;         // plane m + 1
                const int sb = (u - 2 - h + 2 * L) % L;     // plane m - 1
                const uint32_t wys = sWy[sm][gwm][wwm];
                float pl, pr, wxl, wxo, wyo, wym, wzo, wzb, bc;
                if (eact == 0) {
                    bc = RB16 ? h_el<0>(rbh[RB16 ? sm : 0]) : rb[0][RB16 ? 0 : sm];
                    pl = from_left(p[1][sm]);
                    pr = p[1][sm];
                    const float w = bf_el<1>((uint32_t)lane_shr1((int)rwx[sm]));
                    wxl = row_start ? 0.f : w;
                    wxo = bf_el<0>(rwx[sm]);
                    wyo = bf_el<0>(rwy[sm]);
                    wym = bf_el<0>(wys);
                    wzo = bf_el<0>(rwz[sm]);
                    wzb = bf_el<0>(rwz[sb]);
                } else {
                    bc = RB16 ? h_el<1>(rbh[RB16 ? sm : 0]) : rb[1][RB16 ? 0 : sm];
                    pl = p[0][sm];
                    pr = from_right(p[0][sm]);
                    wxl = bf_el<0>(rwx[sm]);
                    wxo = bf_el<1>(rwx[sm]);
                    wyo = bf_el<1>(rwy[sm]);
                    wym = bf_el<1>(wys);
                    wzo = bf_el<1>(rwz[sm]);
                    wzb = bf_el<1>(rwz[sb]);
                }
                float acc = bc;
                acc += wxo * pr;
                acc += wxl * pl;
                acc += wyo * sP[sm][y + 1][cp1[eact]];
                acc += wym * sP[sm][y - 1][cm1[eact]];
                acc += wzo * p[eact][sa];
                acc += wzb * p[eact][sb];
                const float pc = p[eact][sm];
                const float pv = pc + omega * (acc * rinv[eact][sm] - pc);
                p[eact][sm] = pv;
                sP[sm][y][c0[eact]] = pv;
            }
            const int mo = k - R;
            const uint32_t oo = mo >= k0 && mo < k1 ? (uint32_t)mo * pbytes : 0x80000000u;
            const int so = (u - R + 2 * L) % L;
            tse_f2 out;
            out.x = p[0][so];
            out.y = p[1][so];
            buf_st2(out, rOut, oout + oo);
        }
    }
}

template <int NS, int BYT, bool RB16>
__device__ __forceinline__ void
rb_zmarch_tile_x5_in(const ZGeo &g, int tile, const float *__restrict__ Pin,
                     float *__restrict__ Pout, const uint32_t *__restrict__ rec,
                     const void *__restrict__ rhs, float omega, int zero_init,
                     ZRing<float, NS, 2 * BYT> &sP, WyRing<NS, BYT> &sWy) {
    constexpr int R = 2 * NS;
    constexpr int BX = 64;
    constexpr int BY = 2 * BYT;
    constexpr int L = R + 2;
    static_assert(BY - 2 * R >= 1, "tile too small");

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);
    const int ty = zc - zc2 * g.tiles_y;
    const Span sx = tile_span(g.edge, tx, g.tiles_x, g.nx, BX, R);
    const Span sy = tile_span(g.edge, ty, g.tiles_y, g.ny, BY, R);
    const int ip = threadIdx.x & 31;
    const int ly = ((threadIdx.y >> 1) << 2) + (threadIdx.y & 1) + ((threadIdx.x >> 5) << 1);
    const int i = sx.x0 + 2 * ip;
    const int j = sy.x0 + ly;
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const uint32_t plane = (uint32_t)g.nx * (uint32_t)g.ny;
    const uint32_t pbytes = plane * 4u, rbytes = plane * 8u, hbytes = plane * 2u;
    const uint32_t fbytes = pbytes * (uint32_t)g.nz;
    const auto rP = buf_rsrc(Pin, zero_init ? 0u : fbytes);
    const auto rRec = buf_rsrc(rec, rbytes * (uint32_t)g.nz);
    const auto rB = buf_rsrc(rhs, RB16 ? hbytes * (uint32_t)g.nz : fbytes);
    const auto rOut = buf_rsrc(Pout, fbytes);
    const bool in_dom = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
    const uint32_t cell = (uint32_t)(i + g.nx * j);
    const uint32_t ocol = in_dom ? cell * 4u : kBufMax;
    const uint32_t orcol = in_dom ? cell * 8u : kBufMax;
    const uint32_t ohcol = in_dom ? cell * 2u : kBufMax;
    const bool own = in_dom && i >= sx.lo && i < sx.hi && j >= sy.lo && j < sy.hi;
    const uint32_t oout = own ? ocol : kBufMax;
    // every ring slot starts at zero: the phases of the first steps read
    // planes below the chunk, which nothing writes (and the zero guard rows
    // 0 and BY+1 stay zero)
    {
        const int t = threadIdx.x + 64 * threadIdx.y;
        for (int q = t; q < L * (BY + 2) * 64; q += 64 * BYT) (&sP[0][0][0])[q] = 0.f;
        for (int q = t; q < L * (BYT + 1) * 64; q += 64 * BYT) (&sWy[0][0][0])[q] = 0u;
        __syncthreads();
    }
    const int q = __builtin_amdgcn_readfirstlane((j + kl0 - 1) & 1);
    if (q == 0)
        rb_zmarch_x5_march<NS, BYT, 0, RB16>(ly, ip, k0, k1, kl0, kl1, pbytes, rbytes,
                                             hbytes, ocol, orcol, ohcol, oout, rP, rRec,
                                             rB, rOut, omega, sP, sWy);
    else
        rb_zmarch_x5_march<NS, BYT, 1, RB16>(ly, ip, k0, k1, kl0, kl1, pbytes, rbytes,
                                             hbytes, ocol, orcol, ohcol, oout, rP, rRec,
                                             rB, rOut, omega, sP, sWy);
}

template <int NS, int BYT, bool RB16>
__device__ __forceinline__ void
rb_zmarch_tile_x5(const ZGeo &g, int tile, const float *__restrict__ Pin,
                  float *__restrict__ Pout, const uint32_t *__restrict__ rec,
                  const void *__restrict__ rhs, float omega, int zero_init) {
    __shared__ ZRing<float, NS, 2 * BYT> sP;
    __shared__ WyRing<NS, BYT> sWy;
    rb_zmarch_tile_x5_in<NS, BYT, RB16>(g, tile, Pin, Pout, rec, rhs, omega, zero_init,
                                        sP, sWy);
}

struct FastOp {  // natural-layout operator arrays
    double *Wx, *Wy, *Wz, *D, *invD, *b;
    float *Wx32, *Wy32, *Wz32, *invD32;
    uint16_t *rec16;   // variant-5 pair records (bfloat16): Wx, Wy, Wz, V R of
                       // pair q at 8q..8q+7, element 0 first; V R = +inf marks
                       // a Dirichlet cell
};


struct TrOp {  // natural layout, coefficient f of cell t at c[f*ld + t]
    double *c, *D, *invD, *b;
    float *c32, *invD32;
    long N;   // cells
    long ld;  // element distance between the coefficient arrays
};


// One tile of the fused red/black z-march for the non-symmetric transport
// operator (blocks of 64 x BY threads).  Same pipeline as rb_zmarch_tile
// (one barrier per plane step, NS sweeps per launch, 2*NS halo ring,
// ping-pong); the coefficients are the cell's own, so only the unknown goes
// through LDS.
template <typename T, int NS, int BY>
__device__ __forceinline__ void
tr_zmarch_tile(const ZGeo &g, int tile, const T *__restrict__ Pin,
               T *__restrict__ Pout, const T *__restrict__ coef, long ld,
               const T *__restrict__ invD, const T *__restrict__ rhs, T omega,
               int zero_init) {
    constexpr int R = 2 * NS;
    constexpr int BX = 64;
    constexpr int TX = BX - 2 * R;
    constexpr int TY = BY - 2 * R;
    constexpr int NSLOT = R + 2;
    static_assert(TY >= 1, "tile too small");
    __shared__ T sP[NSLOT][BY][BX];

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);
    const int ty = zc - zc2 * g.tiles_y;
    const int lx = threadIdx.x, ly = threadIdx.y;
    const int i = tx * TX - R + lx;
    const int j = ty * TY - R + ly;
    const bool in_dom = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
    const int d = min(min(lx, BX - 1 - lx), min(ly, BY - 1 - ly));
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const long plane = (long)g.nx * g.ny;
    const long col = in_dom ? (long)i + (long)g.nx * j : 0;
    const int par = (i + j) & 1;

    T p[R + 2], rinv[R + 1], rb[R + 1], rc[6][R + 1];
#pragma unroll
    for (int q = 0; q < R + 2; ++q) p[q] = T(0);
#pragma unroll
    for (int q = 0; q < R + 1; ++q) {
        rinv[q] = T(0);
        rb[q] = T(0);
#pragma unroll
        for (int f = 0; f < 6; ++f) rc[f][q] = T(0);
    }
    T np = T(0), ni = T(0), nb = T(0), nc[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) nc[f] = T(0);
    if (in_dom && kl0 <= kl1) {
        const long o = col + plane * kl0;
        np = zero_init ? T(0) : Pin[o];
        ni = invD[o];
        nb = rhs[o];
#pragma unroll
        for (int f = 0; f < 6; ++f) nc[f] = coef[f * ld + o];
    }
    int s0 = kl0 % NSLOT;
    for (int k = kl0; k <= k1 + R - 1; ++k) {
        const T p0 = np, i0 = ni, b0 = nb;
        T c0[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            c0[f] = nc[f];
            nc[f] = T(0);
        }
        np = ni = nb = T(0);
        if (in_dom && k + 1 <= kl1) {
            const long o = col + plane * (k + 1);
            np = zero_init ? T(0) : Pin[o];
            ni = invD[o];
            nb = rhs[o];
#pragma unroll
            for (int f = 0; f < 6; ++f) nc[f] = coef[f * ld + o];
        }
#pragma unroll
        for (int q = R + 1; q > 0; --q) p[q] = p[q - 1];
#pragma unroll
        for (int q = R; q > 0; --q) {
            rinv[q] = rinv[q - 1];
            rb[q] = rb[q - 1];
#pragma unroll
            for (int f = 0; f < 6; ++f) rc[f][q] = rc[f][q - 1];
        }
        p[0] = p0;
        rinv[0] = i0;
        rb[0] = b0;
#pragma unroll
        for (int f = 0; f < 6; ++f) rc[f][0] = c0[f];
        sP[s0][ly][lx] = p0;
        __syncthreads();
#pragma unroll
        for (int h = 0; h < R; ++h) {
            const int m = k - 1 - h;
            int s = s0 - 1 - h;
            if (s < 0) s += NSLOT;
            if (m >= kl0 && m <= kl1 && d >= h + 1 && in_dom &&
                ((par + m) & 1) == (h & 1) && rinv[1 + h] != T(0)) {
                T acc = rb[1 + h];
                acc += rc[0][1 + h] * sP[s][ly][lx + 1];
                acc += rc[1][1 + h] * sP[s][ly][lx - 1];
                acc += rc[2][1 + h] * sP[s][ly + 1][lx];
                acc += rc[3][1 + h] * sP[s][ly - 1][lx];
                acc += rc[4][1 + h] * p[h];
                acc += rc[5][1 + h] * p[h + 2];
                const T pc = p[1 + h];
                const T pn = pc + omega * (acc * rinv[1 + h] - pc);
                p[1 + h] = pn;
                sP[s][ly][lx] = pn;
            }
        }
        const int mo = k - R;
        if (mo >= k0 && mo < k1 && d >= R && in_dom) Pout[col + plane * mo] = p[R];
        s0 = (s0 + 1 == NSLOT) ? 0 : s0 + 1;
    }
}


// Transport z-march with the structure of rb_zmarch_tile_xr: rings of
// length L = 2*NS + 2 rotated by unrolling, the eight coefficient planes
// (six faces, invD, rhs) loaded by buffer loads straight into their slots
// one step ahead, out-of-grid lanes and planes past the tile reading 0, x
// neighbours from the adjacent lanes (DPP), y neighbours through LDS.  Same
// operands in the same order as tr_zmarch_tile: bit-identical results.
template <typename T, int NS, int BY>
__device__ __forceinline__ void
tr_zmarch_tile_xr(const ZGeo &g, int tile, const T *__restrict__ Pin,
                  T *__restrict__ Pout, const T *__restrict__ coef, long ld,
                  const T *__restrict__ invD, const T *__restrict__ rhs, T omega,
                  int zero_init) {
    constexpr int R = 2 * NS;
    constexpr int BX = 64;
    constexpr int TX = BX - 2 * R;
    constexpr int TY = BY - 2 * R;
    constexpr int L = R + 2;
    static_assert(TY >= 1, "tile too small");
    __shared__ T sP[L][BY + 2][BX];   // rows shifted by one, zero guards 0 and BY+1

    const int zc = (int)fdiv((uint32_t)tile, g.div_tx);
    const int tx = tile - zc * g.tiles_x;
    const int zc2 = (int)fdiv((uint32_t)zc, g.div_ty);
    const int ty = zc - zc2 * g.tiles_y;
    const int lx = threadIdx.x, ly = threadIdx.y;
    const Span sx = tile_span(g.edge, tx, g.tiles_x, g.nx, BX, R);
    const Span sy = tile_span(g.edge, ty, g.tiles_y, g.ny, BY, R);
    const int i = sx.x0 + lx;
    const int j = sy.x0 + ly;
    const bool in_dom = (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
    const int d = min(min(sx.lfree ? kFar : lx, sx.hfree ? kFar : BX - 1 - lx),
                      min(sy.lfree ? kFar : ly, sy.hfree ? kFar : BY - 1 - ly));
    const int k0 = zc2 * g.zchunk;
    const int k1 = min(k0 + g.zchunk, g.nz);
    const int kl0 = max(k0 - R, 0);
    const int kl1 = min(k1 + R, g.nz) - 1;
    const int klast = k1 + R - 1;
    const uint32_t pbytes = (uint32_t)g.nx * (uint32_t)g.ny * (uint32_t)sizeof(T);
    const uint32_t fbytes = pbytes * (uint32_t)g.nz;
    const auto rP = buf_rsrc(Pin, zero_init ? 0u : fbytes);
    const auto rI = buf_rsrc(invD, fbytes), rB = buf_rsrc(rhs, fbytes);
    const auto rOut = buf_rsrc(Pout, fbytes);
    __amdgpu_buffer_rsrc_t rC[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) rC[f] = buf_rsrc(coef + f * ld, fbytes);
    bool dm[R];
#pragma unroll
    for (int h = 0; h < R; ++h) dm[h] = in_dom && d >= h + 1;
    const bool cm = ((i + j + kl0 - 1) & 1) == 0;   // colour test, (k-kl0) even
    const uint32_t ocol = in_dom ? (uint32_t)(i + g.nx * j) * (uint32_t)sizeof(T) : kBufMax;
    const bool own = in_dom && i >= sx.lo && i < sx.hi && j >= sy.lo && j < sy.hi;
    const uint32_t oout = own ? ocol : kBufMax;

    T p[L], rinv[L], rb[L], rc[6][L];
#pragma unroll
    for (int r = 1; r < L; ++r) {
        p[r] = rinv[r] = rb[r] = T(0);
#pragma unroll
        for (int f = 0; f < 6; ++f) rc[f][r] = T(0);
    }
    p[0] = T(0);
    T np;
    {
        const uint32_t o = ocol + (kl0 <= kl1 ? (uint32_t)kl0 * pbytes : 0x80000000u);
        np = buf_ld(rP, o, T());
        rinv[0] = buf_ld(rI, o, T());
        rb[0] = buf_ld(rB, o, T());
#pragma unroll
        for (int f = 0; f < 6; ++f) rc[f][0] = buf_ld(rC[f], o, T());
    }
    if (ly == 0)
#pragma unroll
        for (int u = 0; u < L; ++u) sP[u][0][lx] = T(0);
    if (ly == BY - 1)
#pragma unroll
        for (int u = 0; u < L; ++u) sP[u][BY + 1][lx] = T(0);
    for (int kb = kl0; kb <= klast; kb += L) {
#pragma unroll
        for (int u = 0; u < L; ++u) {
            const int k = kb + u;
            if (k > klast) break;
            const int un = (u + 1) % L;
            p[u] = np;
            sP[u][ly + 1][lx] = np;
            const uint32_t o =
                ocol + (k + 1 <= kl1 ? (uint32_t)(k + 1) * pbytes : 0x80000000u);
            np = buf_ld(rP, o, T());
            rinv[un] = buf_ld(rI, o, T());
            rb[un] = buf_ld(rB, o, T());
#pragma unroll
            for (int f = 0; f < 6; ++f) rc[f][un] = buf_ld(rC[f], o, T());
            __syncthreads();
            const bool col_ok = (u & 1) ? !cm : cm;
#pragma unroll
            for (int h = 0; h < R; ++h) {
                const int sm = (u - 1 - h + 2 * L) % L;     // plane m = k-1-h
                const int sa = (u - h + 2 * L) % L;         // plane m + 1
                const int sb = (u - 2 - h + 2 * L) % L;     // plane m - 1
                const T pl = from_left(p[sm]);
                const T pr = from_right(p[sm]);
                if (dm[h] && col_ok && rinv[sm] != T(0)) {
                    T acc = rb[sm];
                    acc += rc[0][sm] * pr;
                    acc += rc[1][sm] * pl;
                    acc += rc[2][sm] * sP[sm][ly + 2][lx];
                    acc += rc[3][sm] * sP[sm][ly][lx];
                    acc += rc[4][sm] * p[sa];
                    acc += rc[5][sm] * p[sb];
                    const T pc = p[sm];
                    const T pn = pc + omega * (acc * rinv[sm] - pc);
                    p[sm] = pn;
                    sP[sm][ly + 1][lx] = pn;
                }
            }
            const int mo = k - R;
            buf_st(p[(u - R + 2 * L) % L], rOut,
                   oout + (mo >= k0 && mo < k1 ? (uint32_t)mo * pbytes : 0x80000000u));
        }
    }
}

__device__ inline double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Residual of one cell t = i + nx*(j + ny*k), accumulated into (sr, sb) =
// (||r||^2, ||b||^2) over the updated cells.
// mode 0: check only, on p_in.
// mode 1: defect correction: p_out = p_in + e; r = b - A p_out written as FP32.
// mode 2: r = b - A p_in written as FP32 (first correction cycle).
template <int MODE>
__device__ __forceinline__ void
p_residual_cell(int nx, int ny, int nz, int i, int j, int k, long t,
                const double *__restrict__ pin, const float *__restrict__ e,
                double *__restrict__ pout, const FastOp &op,
                float *__restrict__ r32, double &sr, double &sb) {
    const long plane = (long)nx * ny;
    auto P = [&](long o) -> double {
        return MODE == 1 ? pin[o] + (double)e[o] : pin[o];
    };
    const double pc = P(t);
    if (MODE == 1) pout[t] = pc;
    double r = 0.0;
    if (op.invD[t] > 0.0) {
        double acc = op.b[t];
        if (i < nx - 1) acc += op.Wx[t] * P(t + 1);
        if (i > 0) acc += op.Wx[t - 1] * P(t - 1);
        if (j < ny - 1) acc += op.Wy[t] * P(t + nx);
        if (j > 0) acc += op.Wy[t - nx] * P(t - nx);
        if (k < nz - 1) acc += op.Wz[t] * P(t + plane);
        if (k > 0) acc += op.Wz[t - plane] * P(t - plane);
        r = acc - op.D[t] * pc;
        sr += r * r;
        sb += op.b[t] * op.b[t];
    }
    if (MODE != 0) r32[t] = (float)r;
}

template <int MODE>
__device__ __forceinline__ void
tr_residual_cell(int nx, int ny, int nz, int i, int j, int k, long t,
                 const double *__restrict__ pin, const float *__restrict__ e,
                 double *__restrict__ pout, const TrOp &op,
                 float *__restrict__ r32, double &sr, double &sb) {
    const long L = op.ld;
    const long plane = (long)nx * ny;
    auto X = [&](long o) -> double {
        return MODE == 1 ? pin[o] + (double)e[o] : pin[o];
    };
    const double xc = X(t);
    if (MODE == 1) pout[t] = xc;
    double r = 0.0;
    if (op.invD[t] != 0.0) {
        double acc = op.b[t];
        if (i < nx - 1) acc += op.c[0 * L + t] * X(t + 1);
        if (i > 0) acc += op.c[1 * L + t] * X(t - 1);
        if (j < ny - 1) acc += op.c[2 * L + t] * X(t + nx);
        if (j > 0) acc += op.c[3 * L + t] * X(t - nx);
        if (k < nz - 1) acc += op.c[4 * L + t] * X(t + plane);
        if (k > 0) acc += op.c[5 * L + t] * X(t - plane);
        r = acc - op.D[t] * xc;
        sr += r * r;
        sb += op.b[t] * op.b[t];
    }
    if (MODE != 0) r32[t] = (float)r;
}

// Block sum of (a, c) written to out[0], out[1] by thread 0 (blocks of kBlock).
__device__ inline void block_sum2_to(double a, double c, double *out) {
    __shared__ double s1[kBlock / 64], s2[kBlock / 64];
    a = wsum(a);
    c = wsum(c);
    if ((threadIdx.x & 63) == 0) {
        s1[threadIdx.x >> 6] = a;
        s2[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0;
        for (int q = 0; q < kBlock / 64; ++q) {
            x += s1[q];
            y += s2[q];
        }
        out[0] = x;
        out[1] = y;
    }
}

enum SysKind { SYS_P = 0, SYS_TR = 1 };

// One linear system in the performance form, owned by the context.
//   pressure:  buf64 = Wx Wy Wz D invD b | p0 p1
//              buf32 = Wx Wy Wz invD rec16(2 arrays) r | e0 e1
//   transport: buf64 = c[6] D invD b | p0 p1          buf32 = c[6] invD r | e0 e1
struct FastSys {
    int kind = SYS_P;
    // row length of the natural layout: nx for transport, nx rounded up to
    // even for pressure (an odd nx gets one inactive column: zero faces,
    // 1/diagonal 0, storage term +inf, iterate 0), so that every pressure
    // system can run the cell-pair sweeps
    int nxf = 0;
    long N = 0;                  // nxf * ny * nz
    long ld = 0;          // element distance between the carved arrays
    double *buf64 = nullptr;
    float *buf32 = nullptr;
    double *partials = nullptr;
    double *red = nullptr;
    double *red_host = nullptr;
    int cur = 0;          // which of p0, p1 holds the iterate
    bool ready = false;

    int n64() const { return kind == SYS_P ? 6 : 9; }
    int n32() const { return kind == SYS_P ? 7 : 8; }
    double *p(int q) const { return buf64 + (n64() + q) * ld; }
    float *r32() const { return buf32 + (n32() - 1) * ld; }
    float *e(int q) const { return buf32 + (n32() + q) * ld; }
    FastOp pop() const {
        FastOp op;
        op.Wx = buf64;
        op.Wy = buf64 + ld;
        op.Wz = buf64 + 2 * ld;
        op.D = buf64 + 3 * ld;
        op.invD = buf64 + 4 * ld;
        op.b = buf64 + 5 * ld;
        op.Wx32 = buf32;
        op.Wy32 = buf32 + ld;
        op.Wz32 = buf32 + 2 * ld;
        op.invD32 = buf32 + 3 * ld;
        op.rec16 = (uint16_t *)(buf32 + 4 * ld);
        return op;
    }
    TrOp trop() const {
        TrOp op;
        op.c = buf64;
        op.D = buf64 + 6 * ld;
        op.invD = buf64 + 7 * ld;
        op.b = buf64 + 8 * ld;
        op.c32 = buf32;
        op.invD32 = buf32 + 6 * ld;
        op.N = N;
        op.ld = ld;
        return op;
    }
};


FastSys *system_of(tse_ctx *ctx, int kind);
int ensure_sys(tse_ctx *ctx, FastSys *s);
// Row length of kind's natural layout on grid G (see FastSys::nxf).
inline int fast_nx(const Geo &G, int kind) { return kind == SYS_P ? (G.nx + 1) & ~1 : G.nx; }
// G with the natural layout's row length, for the sweep and residual tilings
// (which read nx, ny, nz only; never pass it to a split-layout kernel).
inline Geo fast_tiling_geo(const Geo &G, int kind) {
    Geo F = G;
    F.nx = fast_nx(G, kind);
    return F;
}
bool fast_plan_ok(const tse_sweep_plan &pl, int fp32, int kind);
bool fast_plan_fits(const tse_sweep_plan &pl, long N, int fp32);
bool fast_plan_grid_ok(const tse_sweep_plan &pl, int nx);
ZGeo make_zgeo(const Geo &G, int ns, int by, int zchunk, int edge);

}  // namespace tse
