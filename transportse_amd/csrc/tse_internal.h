// tse_internal.h -- shared definitions of the gfx950 TRANSPORTSE core.
//
// Kernels work on the checkerboard-split layout declared in tse_hip.h:
// colour c of cell (i,j,k) is (i+j+k)&1 and its slot is
// c*S + (k*ny + j)*nxh + (i>>1).  A colour half-sweep therefore reads and
// writes dense, x-contiguous arrays: the own colour at the slot, the other
// colour at the same slot (y/z neighbours) or one slot left/right (x).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tse_hip.h"

namespace tse {

// ---- errors ----------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define TSE_HIP_TRY(expr)                                                  \
    do {                                                                   \
        hipError_t _e = (expr);                                            \
        if (_e != hipSuccess)                                              \
            return ::tse::fail(TSE_ERR_HIP, std::string(#expr) + ": " +    \
                                                hipGetErrorString(_e));    \
    } while (0)

// ---- fast unsigned division for n < 2^31 -------------------------------------
// q = (umulhi(n, mul) + n) >> shift with shift = ceil(log2 d) and
// mul = floor(2^32 (2^shift - d) / d) + 1 (round-up reciprocal method).
struct FastDiv {
    uint32_t d, mul, shift;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0u, 0u};
    while ((1ull << f.shift) < d) f.shift++;
    f.mul = (uint32_t)(((1ull << 32) * ((1ull << f.shift) - d)) / d + 1ull);
    return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv &f) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t hi = __umulhi(n, f.mul);
#else
    uint32_t hi = (uint32_t)(((uint64_t)n * f.mul) >> 32);
#endif
    return (hi + n) >> f.shift;
}

// ---- geometry passed by value to kernels -------------------------------------
struct Geo {
    int nx, ny, nz, nxh;
    int S;             // slots per colour = nxh*ny*nz
    int Sp;            // record slots per colour: S rounded up to even
    long Fr;           // record field stride: 2*Sp + kRecPad
    FastDiv div_nxh;   // slot -> row
    FastDiv div_ny;    // row  -> (j, k)
    const double *dx, *dy, *dz;  // device copies
    double g, sinb, cosb;
};

struct Cell {
    int i, j, k, row, ih, s;  // s = c ^ ((j+k)&1): x-offset of the cell in its pair
    bool valid;
};

// Decode slot r (0 <= r < S) of colour c.
__device__ inline Cell decode(const Geo &G, int c, int r) {
    Cell q;
    q.row = (int)fdiv((uint32_t)r, G.div_nxh);
    q.ih = r - q.row * G.nxh;
    q.k = (int)fdiv((uint32_t)q.row, G.div_ny);
    q.j = q.row - q.k * G.ny;
    q.s = c ^ ((q.j + q.k) & 1);
    q.i = 2 * q.ih + q.s;
    q.valid = q.i < G.nx;
    return q;
}

// Slot of an arbitrary cell.
__host__ __device__ inline int slot(const Geo &G, int i, int j, int k) {
    return ((i + j + k) & 1) * G.S + (k * G.ny + j) * G.nxh + (i >> 1);
}

// Coefficient records and masks of the faithful solvers use their own
// layout: colour c starts at c*Sp (Sp even, so a pair of slots (2m, 2m+1) is
// one 16-byte load), and field f of a record starts at f*Fr.  Fr = 2*Sp +
// kRecPad: with a field stride that is a large power of two the 13 record
// streams of a sweep fall on the same HBM channels at the same time (a
// 512x512x256 half-sweep ran at 52 % of the HBM peak, 67 % with the pad).
// A pad slot (r = S when S is odd) holds a zero record and mask 0.
constexpr long kRecPad = 544;
__host__ __device__ inline long rec_slot(const Geo &G, int i, int j, int k) {
    return (long)((i + j + k) & 1) * G.Sp + (long)(k * G.ny + j) * G.nxh + (i >> 1);
}

// ---- device control block of one SOR solve ------------------------------------
// Written by the control kernel after every iteration; the sweep kernels read
// `status` and return at once when the solve has stopped, so the host can
// queue iterations ahead without a round trip per iteration.
enum CtlStatus { RUNNING = 0, CONVERGED = 1, NEED_HALVE = 2, FAILED = 3 };
enum CheckMode { CHECK_P_SCHEDULE = 0, CHECK_EVERY = 1 };

struct Ctl {
    int32_t num_it, num_sub_it, status, check_mode;
    int32_t max_it, omega_zero, pad0, pad1;
    double t, dt, diff, conv, inc_it, max_dt, min_dt, pad2;
};

__host__ __device__ inline bool is_check_iter(int mode, int it) {
    return mode == CHECK_EVERY || it < 50 || (it % 10) == 0;
}

// ---- launch shapes ------------------------------------------------------------------
constexpr int kBlock = 256;          // 4 wavefronts of 64
// Per-cell coefficient records of the faithful solvers (tse_faithful.hip):
// compute_P: A_, C_, E_, G_, I_, K_, A_B_, C_D_, I_J_, K_L_, R*Pn, WP, R+M;
// transport: a(x+), a(x-), a(y+), a(y-), a(z+), a(z-), X, R*Xn, W, E,
// R - omega*M.  Each field is a 2*S array in the split layout.
constexpr int kPRec = 13;
constexpr int kTRec = 11;
constexpr int kMaxSweepBlocks = 65536;
constexpr int kCtlBlock = 1024;      // control kernel: reads 2 x sweep blocks partials
// Grids of at most kFuseBlocks sweep blocks run the control step inside the
// next red half-sweep (one partial pair per thread of a kBlock block).
constexpr int kFuseBlocks = kBlock / 2;
// Doubles of one iteration's norm partials: [2 colours][kMaxSweepBlocks][2].
constexpr long kPartialsPerIter = 2L * 2 * kMaxSweepBlocks;

// Control wiring of one faithful half-sweep launch.  `in` is the solver
// state the sweep runs under (null: fixed sweeps, no control); with `fuse`
// the sweep first applies the previous iteration's control step to `in`
// with its partials `prev`, and block 0 stores the result at `out`.
struct SweepCtl {
    const Ctl *in;
    Ctl *out;
    const double *prev;
    double *partials;
    int nblocks;
    int fuse;
};

inline int sweep_blocks(int S) {
    long b = ((long)S + kBlock - 1) / kBlock;
    if (b > kMaxSweepBlocks) b = kMaxSweepBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

// Blocks of a faithful half-sweep: one pair of slots per thread, up to
// kMaxSweepBlocks (beyond that the blocks stride over the slots).
inline int pair_sweep_blocks(const Geo &G) { return sweep_blocks(G.Sp / 2); }

// Record buffer sizes (doubles / bytes) for nfields fields.
inline size_t rec_buffer_elems(const Geo &G, int nfields) { return (size_t)nfields * G.Fr; }
inline size_t mask_bytes(const Geo &G) { return 2 * (size_t)G.Sp; }

}  // namespace tse

// ---- context ----------------------------------------------------------------------
struct tse_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool grid_set = false;
    tse::Geo geo{};
    double *d_dxyz = nullptr;  // dx | dy | dz
    // scratch owned by the context (grown on demand)
    double *rec = nullptr;
    size_t rec_elems = 0;
    uint8_t *mask = nullptr;
    size_t mask_elems = 0;
    double *partials = nullptr;  // [2 iteration parities][kPartialsPerIter] (4 MiB)
    tse::Ctl *ctl = nullptr;      // device, two slots (see sor_solve)
    tse::Ctl *ctl_host = nullptr; // pinned host mirror
    double *red_out = nullptr;    // device scratch (4 doubles) for reductions
    double *red_host = nullptr;   // pinned
    void *fast = nullptr;         // performance-form pressure system (tse_fast.hip)
    void *fast_tr = nullptr;      // performance-form transport system
};

namespace tse {
int ensure_rec(tse_ctx *ctx, size_t elems);
int ensure_mask(tse_ctx *ctx, size_t elems);
int check_ctx(tse_ctx *ctx);

// launchers implemented in tse_faithful.hip
int launch_p_assemble(tse_ctx *ctx, const tse_p_fields *f, double dt);
int launch_p_sweep(tse_ctx *ctx, double *P, int colour, double sor,
                   const SweepCtl &sc);
int launch_control(tse_ctx *ctx, const Ctl *in, const double *partials,
                   int nblocks);
int launch_q(tse_ctx *ctx, const double *P, const double *BP,
             const double *Kx, const double *Ky, const double *Kz,
             double *qx, double *qy, double *qz, const double *nue,
             const double *rho_f);
int launch_flow_error(tse_ctx *ctx, const double *P, const double *Pn,
                      const double *qx, const double *qy, const double *qz,
                      const double *WP, const double *phi, const double *c_f,
                      const double *c_r, const double *rho_f, double dt,
                      double *e);
int launch_c_assemble(tse_ctx *ctx, const tse_c_fields *f, double dt,
                      double omega);
int launch_t_assemble(tse_ctx *ctx, const tse_t_fields *f, double dt,
                      double omega);
int launch_tr_sweep(tse_ctx *ctx, double *X, int colour, double sor,
                    double omega, const SweepCtl &sc);
int launch_l2(tse_ctx *ctx, const double *New, const double *Old);
int launch_smolarkiewicz(tse_ctx *ctx, const double *C, const double *qx,
                         const double *qy, const double *qz, double *vx,
                         double *vy, double *vz, double dt, double eps,
                         double Sc);
// implemented in tse_fast.hip
void fast_free(tse_ctx *ctx);
// compute_P / compute_C / compute_T (kind 0 / 1 / 2) in a performance form,
// with the reference's time-step control around each attempt
int fast_compute(tse_ctx *ctx, int kind, const void *fields,
                 const tse_sor_params *prm, tse_result *out);
}  // namespace tse
