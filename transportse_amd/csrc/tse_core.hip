// tse_core.hip -- context, layout conversion and the C-ABI entry points.
//
// The SOR solvers run as a queue of (red sweep, black sweep, control) launch
// triples on the context's stream (on small grids the control step runs
// inside the next red sweep, see sor_solve).  The control step keeps the
// solver state (iteration counters, t, dt, status) in device memory and the
// sweeps return immediately once the solve has stopped, so the host enqueues
// iterations in growing chunks and synchronises once per chunk, not once per
// iteration.
// Only a dt halving needs the host: it restores the field, halves dt,
// re-assembles the coefficient records (the storage term depends on dt) and
// resumes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include "tse_internal.h"

namespace tse {

// ---- fixed scratch of tse_compute_p_noalloc -------------------------------------------------
// The coefficient records and the cell mask of the faithful compute_P,
// which tse_compute_p keeps in context buffers grown with hipMalloc, live
// for tse_compute_p_noalloc in this file-scope device array: each GPU the
// library runs on holds one copy, at its own address, until the process
// ends.  Host code never uses the array's name as an address: tse_ctx_create
// looks up the device's copy once per device with hipGetSymbolAddress (with
// that device current) and keeps it in g_noalloc_ptr.
constexpr size_t kNoallocScratchBytes = 64ull << 20;
constexpr int kNoallocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
static void *g_noalloc_ptr[kNoallocMaxDevices];
static std::mutex g_noalloc_mu;

// Called with `device` current.  A failed lookup leaves the pointer null:
// the context still works, and tse_compute_p_noalloc reports the error.
static void noalloc_init(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p)
        g_noalloc_ptr[device] = p;
    else
        (void)hipGetLastError();
}

static void *noalloc_ptr(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

// Bytes of fixed scratch a grid needs: records, then the mask at the next
// 256-byte boundary.
static size_t noalloc_bytes(const Geo &G, size_t *mask_off) {
    const size_t rec = rec_buffer_elems(G, kPRec) * sizeof(double);
    const size_t off = (rec + 255) / 256 * 256;
    *mask_off = off;
    return off + mask_bytes(G);
}

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int check_ctx(tse_ctx *ctx) {
    if (!ctx) return fail(TSE_ERR_ARG, "null context");
    if (!ctx->grid_set) return fail(TSE_ERR_STATE, "grid not set");
    TSE_HIP_TRY(hipSetDevice(ctx->device));
    return 0;
}

int ensure_rec(tse_ctx *ctx, size_t elems) {
    if (ctx->rec_elems >= elems) return 0;
    if (ctx->rec) {
        TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
        TSE_HIP_TRY(hipFree(ctx->rec));
        ctx->rec = nullptr;
        ctx->rec_elems = 0;
    }
    if (hipMalloc((void **)&ctx->rec, elems * sizeof(double)) != hipSuccess)
        return fail(TSE_ERR_NOMEM, "hipMalloc of coefficient records failed");
    ctx->rec_elems = elems;
    return 0;
}

int ensure_mask(tse_ctx *ctx, size_t elems) {
    if (ctx->mask_elems >= elems) return 0;
    if (ctx->mask) {
        TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
        TSE_HIP_TRY(hipFree(ctx->mask));
        ctx->mask = nullptr;
        ctx->mask_elems = 0;
    }
    if (hipMalloc((void **)&ctx->mask, elems) != hipSuccess)
        return fail(TSE_ERR_NOMEM, "hipMalloc of cell mask failed");
    ctx->mask_elems = elems;
    return 0;
}

// ---- layout conversion ---------------------------------------------------------------------

// Host layout (element strides si, sj, sk; C order has k fastest) <-> split
// layout (x fastest) as a transpose through LDS: a block takes a 64 (i) x 64
// (k) tile of one y row j, so host-side accesses run along k and split-side
// accesses along i (two colour streams of 32 consecutive slots per
// wavefront).  i runs to 2*nxh: for odd nx, cell i = nx stands for the pad
// slot of its colour, which holds 0.
constexpr int kTT = 64;

// Block b takes row j = b / (nti * ntk) and tile t = b % (nti * ntk) of
// nti x ntk tiles (i major).
__device__ inline void tile_origin(const Geo &G, int &i0, int &k0, int &j) {
    const int nti = (2 * G.nxh + kTT - 1) / kTT, ntk = (G.nz + kTT - 1) / kTT;
    const int b = blockIdx.x, t = b % (nti * ntk);
    j = b / (nti * ntk);
    i0 = (t / ntk) * kTT;
    k0 = (t % ntk) * kTT;
}

__global__ void __launch_bounds__(256)
to_split_kernel(Geo G, double *__restrict__ dst, const double *__restrict__ src,
                long si, long sj, long sk) {
    __shared__ double tile[kTT][kTT + 1];
    int i0, k0, j;
    tile_origin(G, i0, k0, j);
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int ii = ty; ii < kTT; ii += 4) {
        const int i = i0 + ii, k = k0 + tx;
        if (i < G.nx && k < G.nz) tile[ii][tx] = src[i * si + j * sj + k * sk];
    }
    __syncthreads();
    const int i = i0 + tx;
    if (i < 2 * G.nxh) {
        for (int kk = ty; kk < kTT && k0 + kk < G.nz; kk += 4)
            dst[slot(G, i, j, k0 + kk)] = i < G.nx ? tile[tx][kk] : 0.0;
    }
}

__global__ void __launch_bounds__(256)
from_split_kernel(Geo G, double *__restrict__ dst, long si, long sj, long sk,
                  const double *__restrict__ src) {
    __shared__ double tile[kTT][kTT + 1];
    int i0, k0, j;
    tile_origin(G, i0, k0, j);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx;
    if (i < G.nx) {
        for (int kk = ty; kk < kTT && k0 + kk < G.nz; kk += 4)
            tile[tx][kk] = src[slot(G, i, j, k0 + kk)];
    }
    __syncthreads();
    for (int ii = ty; ii < kTT; ii += 4) {
        const int ic = i0 + ii, k = k0 + tx;
        if (ic < G.nx && k < G.nz) dst[ic * si + j * sj + k * sk] = tile[ii][tx];
    }
}

// Small fields (a few tiles, mostly empty): one thread per slot instead.
constexpr long kFlatPermSlots = 1L << 20;

__global__ void __launch_bounds__(kBlock)
to_split_flat_kernel(Geo G, double *__restrict__ dst, const double *__restrict__ src,
                     long si, long sj, long sk) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        dst[t] = q.valid ? src[q.i * si + q.j * sj + q.k * sk] : 0.0;
    }
}

__global__ void __launch_bounds__(kBlock)
from_split_flat_kernel(Geo G, double *__restrict__ dst, long si, long sj, long sk,
                       const double *__restrict__ src) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        if (q.valid) dst[q.i * si + q.j * sj + q.k * sk] = src[t];
    }
}

// Launch shape of the two permutation kernels.
static inline dim3 perm_grid(const Geo &G) {
    const long nti = (2L * G.nxh + kTT - 1) / kTT, ntk = (G.nz + kTT - 1) / kTT;
    return dim3((unsigned)(nti * ntk * G.ny));
}


// ---- SOR driver loop shared by compute_P / compute_C / compute_T --------------------------

struct SolveHooks {
    // re-evaluate the coefficient records for time step dt
    int (*assemble)(tse_ctx *, const void *, double, double);
    const void *fields;
    double *X;   // unknown (in/out)
    double *Xn;  // copy at entry
    bool transport;
};

static int p_assemble_hook(tse_ctx *ctx, const void *f, double dt, double) {
    return launch_p_assemble(ctx, (const tse_p_fields *)f, dt);
}
static int c_assemble_hook(tse_ctx *ctx, const void *f, double dt,
                           double omega) {
    return launch_c_assemble(ctx, (const tse_c_fields *)f, dt, omega);
}
static int t_assemble_hook(tse_ctx *ctx, const void *f, double dt,
                           double omega) {
    return launch_t_assemble(ctx, (const tse_t_fields *)f, dt, omega);
}

// Both state slots get the host's state.
static int write_ctl(tse_ctx *ctx, const Ctl &h) {
    *ctx->ctl_host = h;
    for (int k = 0; k < 2; ++k)
        TSE_HIP_TRY(hipMemcpyAsync(ctx->ctl + k, ctx->ctl_host, sizeof(Ctl),
                                   hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

static int sor_solve(tse_ctx *ctx, const SolveHooks &hk,
                     const tse_sor_params *prm, tse_result *out) {
    const Geo &G = ctx->geo;
    const size_t bytes = 2ull * G.S * sizeof(double);
    if (prm->max_it < 1) return fail(TSE_ERR_ARG, "max_it must be >= 1");
    if (!(prm->dt > 0.0)) return fail(TSE_ERR_ARG, "dt must be > 0");
    TSE_HIP_TRY(hipMemcpyAsync(hk.Xn, hk.X, bytes, hipMemcpyDeviceToDevice,
                               ctx->stream));
    Ctl h{};
    h.num_it = 0;
    h.num_sub_it = 0;
    h.status = RUNNING;
    h.check_mode = hk.transport ? CHECK_EVERY : CHECK_P_SCHEDULE;
    h.max_it = prm->max_it;
    h.omega_zero = (hk.transport && prm->omega == 0.0) ? 1 : 0;
    h.t = 0.0;
    h.dt = prm->dt;
    h.diff = 1e100;
    h.conv = prm->conv;
    h.inc_it = prm->inc_it;
    h.max_dt = prm->max_dt;
    h.min_dt = prm->min_dt;
    int rc = write_ctl(ctx, h);
    if (rc) return rc;
    rc = hk.assemble(ctx, hk.fields, h.dt, prm->omega);
    if (rc) return rc;
    // The solver state lives in two device slots.  Iteration group q (red
    // sweep, black sweep) writes its norm partials to parity q & 1.  On
    // larger grids a control launch follows every group and writes both
    // slots.  On grids of at most kFuseBlocks sweep blocks the control step
    // of group q-1 runs inside the red sweep of group q: it reads slot q & 1
    // (the state after group q-2) and stores slot (q-1) & 1, which the black
    // sweep then reads; a control launch ends each chunk of groups and
    // writes both slots again, so the first group of a chunk and the host
    // see one state.
    const int nb = pair_sweep_blocks(G);
    const bool fuse = nb <= kFuseBlocks;
    int halvings = 0;
    int chunk = 4;
    long q = 0;
    for (;;) {
        for (int n = 0; n < chunk; ++n) {
            ++q;
            double *part = ctx->partials + (q & 1) * kPartialsPerIter;
            const double *prev = ctx->partials + ((q - 1) & 1) * kPartialsPerIter;
            Ctl *const before = ctx->ctl + ((q - 1) & 1);  // state after group q-1
            for (int colour = 0; colour < 2; ++colour) {
                SweepCtl sc{before, nullptr, nullptr, part, nb, 0};
                if (colour == 0 && fuse && n > 0)
                    sc = SweepCtl{ctx->ctl + (q & 1), before, prev, part, nb, 1};
                rc = hk.transport
                         ? launch_tr_sweep(ctx, hk.X, colour, prm->sor, prm->omega, sc)
                         : launch_p_sweep(ctx, hk.X, colour, prm->sor, sc);
                if (rc) return rc;
            }
            if (!fuse || n == chunk - 1) {
                rc = launch_control(ctx, before, part, nb);
                if (rc) return rc;
            }
        }
        TSE_HIP_TRY(hipMemcpyAsync(ctx->ctl_host, ctx->ctl, sizeof(Ctl),
                                   hipMemcpyDeviceToHost, ctx->stream));
        TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
        h = *ctx->ctl_host;
        if (h.status == RUNNING) {
            chunk = std::min(chunk * 2, 256);
            continue;
        }
        if (h.status == NEED_HALVE) {
            // tse.py:733-740: dt /= 2, num_sub_it = 0, X = Xn, stop if dt <= min_dt
            h.dt = h.dt / 2;
            h.num_sub_it = 0;
            halvings += 1;
            TSE_HIP_TRY(hipMemcpyAsync(hk.X, hk.Xn, bytes,
                                       hipMemcpyDeviceToDevice, ctx->stream));
            if (h.dt <= h.min_dt) {
                h.status = FAILED;
                break;
            }
            h.status = RUNNING;
            rc = write_ctl(ctx, h);
            if (rc) return rc;
            rc = hk.assemble(ctx, hk.fields, h.dt, prm->omega);
            if (rc) return rc;
            chunk = 4;
            continue;
        }
        break;  // CONVERGED or FAILED
    }
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    out->t = h.t;
    out->dt = h.dt;
    out->converged = h.status == CONVERGED ? 1 : 0;
    out->num_it = h.num_it;
    out->num_halvings = halvings;
    out->status = h.status;
    out->last_diff = h.diff;
    return 0;
}

}  // namespace tse

using namespace tse;

extern "C" {

int tse_abi_version(void) { return TSE_ABI_VERSION; }

const char *tse_last_error(void) { return tse::g_err.c_str(); }

int tse_ctx_create(tse_ctx **out, int device, void *stream) {
    if (!out) return fail(TSE_ERR_ARG, "null out pointer");
    *out = nullptr;
    int ndev = 0;
    TSE_HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(TSE_ERR_ARG, "device index out of range");
    TSE_HIP_TRY(hipSetDevice(device));
    noalloc_init(device);
    tse_ctx *ctx = new tse_ctx();
    ctx->device = device;
    ctx->stream = (hipStream_t)stream;
    if (hipMalloc((void **)&ctx->ctl, 2 * sizeof(Ctl)) != hipSuccess ||
        hipHostMalloc((void **)&ctx->ctl_host, sizeof(Ctl)) != hipSuccess ||
        hipMalloc((void **)&ctx->partials,
                  sizeof(double) * 2 * kPartialsPerIter) != hipSuccess ||
        hipMalloc((void **)&ctx->red_out, 4 * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void **)&ctx->red_host, 4 * sizeof(double)) !=
            hipSuccess) {
        tse_ctx_destroy(ctx);
        return fail(TSE_ERR_NOMEM, "context allocation failed");
    }
    *out = ctx;
    return 0;
}

int tse_ctx_destroy(tse_ctx *ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_dxyz) (void)hipFree(ctx->d_dxyz);
    if (ctx->rec) (void)hipFree(ctx->rec);
    if (ctx->mask) (void)hipFree(ctx->mask);
    if (ctx->partials) (void)hipFree(ctx->partials);
    if (ctx->ctl) (void)hipFree(ctx->ctl);
    if (ctx->ctl_host) (void)hipHostFree(ctx->ctl_host);
    if (ctx->red_out) (void)hipFree(ctx->red_out);
    if (ctx->red_host) (void)hipHostFree(ctx->red_host);
    tse::fast_free(ctx);
    delete ctx;
    return 0;
}

int tse_ctx_set_stream(tse_ctx *ctx, void *stream) {
    if (!ctx) return fail(TSE_ERR_ARG, "null context");
    ctx->stream = (hipStream_t)stream;
    return 0;
}

// ---- graphs of a context's work ------------------------------------------------
struct tse_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int tse_graph_begin(tse_ctx *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->stream) return fail(TSE_ERR_STATE, "graph capture needs a non-null stream");
    TSE_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    return 0;
}

int tse_graph_end(tse_ctx *ctx, tse_graph **out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) return fail(TSE_ERR_ARG, "null out");
    hipGraph_t g = nullptr;
    TSE_HIP_TRY(hipStreamEndCapture(ctx->stream, &g));
    tse_graph *h = new tse_graph();
    h->graph = g;
    if (hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete h;
        return fail(TSE_ERR_HIP, "hipGraphInstantiate failed");
    }
    *out = h;
    return 0;
}

int tse_graph_launch(tse_graph *g, tse_ctx *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!g || !g->exec) return fail(TSE_ERR_ARG, "null graph");
    TSE_HIP_TRY(hipGraphLaunch(g->exec, ctx->stream));
    return 0;
}

int tse_graph_destroy(tse_graph *g) {
    if (!g) return 0;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return 0;
}

int tse_ctx_synchronize(tse_ctx *ctx) {
    if (!ctx) return fail(TSE_ERR_ARG, "null context");
    TSE_HIP_TRY(hipSetDevice(ctx->device));
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tse_set_grid(tse_ctx *ctx, int32_t nx, int32_t ny, int32_t nz,
                 const double *dx, const double *dy, const double *dz,
                 double g, double sin_beta, double cos_beta) {
    if (!ctx) return fail(TSE_ERR_ARG, "null context");
    if (nx < 1 || ny < 1 || nz < 1 || !dx || !dy || !dz)
        return fail(TSE_ERR_ARG, "invalid grid");
    const long nxh = (nx + 1) / 2;
    const long S = nxh * (long)ny * (long)nz;
    if (2 * S >= (1L << 31))
        return fail(TSE_ERR_ARG, "grid too large for 32-bit slot indices");
    TSE_HIP_TRY(hipSetDevice(ctx->device));
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->d_dxyz) {
        TSE_HIP_TRY(hipFree(ctx->d_dxyz));
        ctx->d_dxyz = nullptr;
    }
    const size_t n = (size_t)nx + ny + nz;
    if (hipMalloc((void **)&ctx->d_dxyz, n * sizeof(double)) != hipSuccess)
        return fail(TSE_ERR_NOMEM, "hipMalloc of spacing arrays failed");
    TSE_HIP_TRY(hipMemcpy(ctx->d_dxyz, dx, nx * sizeof(double),
                          hipMemcpyHostToDevice));
    TSE_HIP_TRY(hipMemcpy(ctx->d_dxyz + nx, dy, ny * sizeof(double),
                          hipMemcpyHostToDevice));
    TSE_HIP_TRY(hipMemcpy(ctx->d_dxyz + nx + ny, dz, nz * sizeof(double),
                          hipMemcpyHostToDevice));
    Geo &G = ctx->geo;
    G.nx = nx;
    G.ny = ny;
    G.nz = nz;
    G.nxh = (int)nxh;
    G.S = (int)S;
    G.Sp = (int)(S + (S & 1));
    G.Fr = 2L * G.Sp + kRecPad;
    G.div_nxh = make_fastdiv((uint32_t)nxh);
    G.div_ny = make_fastdiv((uint32_t)ny);
    G.dx = ctx->d_dxyz;
    G.dy = ctx->d_dxyz + nx;
    G.dz = ctx->d_dxyz + nx + ny;
    G.g = g;
    G.sinb = sin_beta;
    G.cosb = cos_beta;
    ctx->grid_set = true;
    return 0;
}

int64_t tse_field_elems(const tse_ctx *ctx) {
    if (!ctx || !ctx->grid_set) return -1;
    return 2LL * ctx->geo.S;
}

int tse_to_split(tse_ctx *ctx, double *dst_split, const double *src,
                 int64_t si, int64_t sj, int64_t sk) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!dst_split || !src) return fail(TSE_ERR_ARG, "null pointer");
    const long F = 2L * ctx->geo.S;
    if (F < kFlatPermSlots)
        hipLaunchKernelGGL(to_split_flat_kernel, dim3((F + kBlock - 1) / kBlock),
                           dim3(kBlock), 0, ctx->stream, ctx->geo, dst_split, src,
                           (long)si, (long)sj, (long)sk);
    else
        hipLaunchKernelGGL(to_split_kernel, perm_grid(ctx->geo), dim3(kTT, 4), 0,
                           ctx->stream, ctx->geo, dst_split, src, (long)si,
                           (long)sj, (long)sk);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int tse_from_split(tse_ctx *ctx, double *dst, int64_t si, int64_t sj,
                   int64_t sk, const double *src_split) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!dst || !src_split) return fail(TSE_ERR_ARG, "null pointer");
    const long F = 2L * ctx->geo.S;
    if (F < kFlatPermSlots)
        hipLaunchKernelGGL(from_split_flat_kernel, dim3((F + kBlock - 1) / kBlock),
                           dim3(kBlock), 0, ctx->stream, ctx->geo, dst, (long)si,
                           (long)sj, (long)sk, src_split);
    else
        hipLaunchKernelGGL(from_split_kernel, perm_grid(ctx->geo), dim3(kTT, 4), 0,
                           ctx->stream, ctx->geo, dst, (long)si, (long)sj,
                           (long)sk, src_split);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int tse_compute_p(tse_ctx *ctx, const tse_p_fields *f,
                  const tse_sor_params *prm, tse_result *out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !prm || !out) return fail(TSE_ERR_ARG, "null argument");
    if (!f->P || !f->Pn || !f->c_f || !f->nue || !f->rho_f || !f->phi ||
        !f->Kx || !f->Ky || !f->Kz || !f->BP || !f->WP || !f->c_r)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (prm->form != TSE_FORM_FAITHFUL64) return fast_compute(ctx, 0, f, prm, out);
    SolveHooks hk{p_assemble_hook, f, f->P, f->Pn, false};
    return sor_solve(ctx, hk, prm, out);
}

int tse_compute_p_noalloc(tse_ctx *ctx, const tse_p_fields *f,
                          const tse_sor_params *prm, tse_result *out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !prm || !out) return fail(TSE_ERR_ARG, "null argument");
    if (!f->P || !f->Pn || !f->c_f || !f->nue || !f->rho_f || !f->phi ||
        !f->Kx || !f->Ky || !f->Kz || !f->BP || !f->WP || !f->c_r)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (prm->form != TSE_FORM_FAITHFUL64)
        return fail(TSE_ERR_ARG, "tse_compute_p_noalloc: only TSE_FORM_FAITHFUL64");
    size_t mask_off = 0;
    const size_t need = noalloc_bytes(ctx->geo, &mask_off);
    if (need > kNoallocScratchBytes)
        return fail(TSE_ERR_ARG,
                    "tse_compute_p_noalloc: the grid needs " + std::to_string(need) +
                        " bytes of scratch, the fixed limit is " +
                        std::to_string(kNoallocScratchBytes) +
                        " (about 6.4e5 cells); use tse_compute_p");
    unsigned char *base = (unsigned char *)noalloc_ptr(ctx->device);
    if (!base)
        return fail(TSE_ERR_STATE,
                    "tse_compute_p_noalloc: no scratch address for this device");
    // The solve runs on the context's record and mask slots pointed at the
    // fixed scratch, sized so that the on-demand growth never triggers; the
    // context's own buffers come back afterwards, whatever the outcome.
    struct Swap {
        tse_ctx *c;
        double *rec;
        size_t rec_elems;
        uint8_t *mask;
        size_t mask_elems;
        ~Swap() {
            c->rec = rec;
            c->rec_elems = rec_elems;
            c->mask = mask;
            c->mask_elems = mask_elems;
        }
    } keep{ctx, ctx->rec, ctx->rec_elems, ctx->mask, ctx->mask_elems};
    ctx->rec = (double *)base;
    ctx->rec_elems = rec_buffer_elems(ctx->geo, kPRec);
    ctx->mask = (uint8_t *)(base + mask_off);
    ctx->mask_elems = mask_bytes(ctx->geo);
    SolveHooks hk{p_assemble_hook, f, f->P, f->Pn, false};
    return sor_solve(ctx, hk, prm, out);
}

int tse_compute_c(tse_ctx *ctx, const tse_c_fields *f,
                  const tse_sor_params *prm, tse_result *out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !prm || !out) return fail(TSE_ERR_ARG, "null argument");
    if (!f->C || !f->Cn || !f->D_fx || !f->D_fy || !f->D_fz || !f->phi ||
        !f->rho_f || !f->qx || !f->qy || !f->qz || !f->BC || !f->WC ||
        !f->RC || !f->e || !f->lf_l || !f->lf_t)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (prm->form != TSE_FORM_FAITHFUL64) return fast_compute(ctx, 1, f, prm, out);
    SolveHooks hk{c_assemble_hook, f, f->C, f->Cn, true};
    return sor_solve(ctx, hk, prm, out);
}

int tse_compute_t(tse_ctx *ctx, const tse_t_fields *f,
                  const tse_sor_params *prm, tse_result *out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !prm || !out) return fail(TSE_ERR_ARG, "null argument");
    if (!f->T || !f->Tn || !f->l_f || !f->cp_f || !f->rho_f || !f->phi ||
        !f->qx || !f->qy || !f->qz || !f->BT || !f->WT || !f->e || !f->l_r ||
        !f->cp_r || !f->rho_r || !f->lt_l || !f->lt_t)
        return fail(TSE_ERR_ARG, "null field pointer");
    if (prm->form != TSE_FORM_FAITHFUL64) return fast_compute(ctx, 2, f, prm, out);
    SolveHooks hk{t_assemble_hook, f, f->T, f->Tn, true};
    return sor_solve(ctx, hk, prm, out);
}

int tse_smolarkiewicz(tse_ctx *ctx, const double *C, const double *qx,
                      const double *qy, const double *qz, double *vx, double *vy,
                      double *vz, double dt, double eps, double Sc) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!C || !qx || !qy || !qz || !vx || !vy || !vz)
        return fail(TSE_ERR_ARG, "null field pointer");
    return launch_smolarkiewicz(ctx, C, qx, qy, qz, vx, vy, vz, dt, eps, Sc);
}

int tse_compute_q(tse_ctx *ctx, const double *P, const double *BP,
                  const double *Kx, const double *Ky, const double *Kz,
                  double *qx, double *qy, double *qz, const double *nue,
                  const double *rho_f) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!P || !BP || !Kx || !Ky || !Kz || !qx || !qy || !qz || !nue || !rho_f)
        return fail(TSE_ERR_ARG, "null field pointer");
    return launch_q(ctx, P, BP, Kx, Ky, Kz, qx, qy, qz, nue, rho_f);
}

int tse_flow_error(tse_ctx *ctx, const double *P, const double *Pn,
                   const double *qx, const double *qy, const double *qz,
                   const double *WP, const double *phi, const double *c_f,
                   const double *c_r, const double *rho_f, double dt_P,
                   double *e) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!P || !Pn || !qx || !qy || !qz || !WP || !phi || !c_f || !c_r ||
        !rho_f || !e)
        return fail(TSE_ERR_ARG, "null field pointer");
    return launch_flow_error(ctx, P, Pn, qx, qy, qz, WP, phi, c_f, c_r, rho_f,
                             dt_P, e);
}

int tse_l2_norm_async(tse_ctx *ctx, const double *New, const double *Old) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!New || !Old) return fail(TSE_ERR_ARG, "null pointer");
    rc = launch_l2(ctx, New, Old);
    if (rc) return rc;
    TSE_HIP_TRY(hipMemcpyAsync(ctx->red_host, ctx->red_out, 2 * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    return 0;
}

int tse_reduction_result(tse_ctx *ctx, double *out, int32_t n) {
    if (!ctx) return fail(TSE_ERR_ARG, "null context");
    if (!out || n < 1 || n > 3) return fail(TSE_ERR_ARG, "n must be 1..3");
    TSE_HIP_TRY(hipSetDevice(ctx->device));
    TSE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < n; ++c) out[c] = ctx->red_host[c];
    return 0;
}

int tse_l2_norm(tse_ctx *ctx, const double *New, const double *Old,
                double *out) {
    if (!out) return fail(TSE_ERR_ARG, "null pointer");
    const int rc = tse_l2_norm_async(ctx, New, Old);
    return rc ? rc : tse_reduction_result(ctx, out, 1);
}

double tse_host_l2_norm_seq(const double *New, const double *Old, int64_t n,
                            int32_t *is_nan) {
    // the two sums are independent: each keeps its sequential order, and on
    // large arrays the second runs on a thread of its own
    double s1 = 0.0, s2 = 0.0;
    auto sum2 = [&]() {
        for (int64_t p = 0; p < n; ++p) s2 += Old[p] * Old[p];
    };
    std::thread other;
    const bool split = n >= (1 << 20);
    if (split) other = std::thread(sum2);
    for (int64_t p = 0; p < n; ++p) {
        const double d = New[p] - Old[p];
        s1 += d * d;
    }
    if (split)
        other.join();
    else
        sum2();
    const double l2 = std::sqrt(s1) / std::sqrt(s2 + 1e-200);
    if (std::isnan(l2)) {
        if (is_nan) *is_nan = 1;
        return 1.0;
    }
    if (is_nan) *is_nan = 0;
    return l2;
}

int tse_p_sweeps(tse_ctx *ctx, const tse_p_fields *f, double dt, double sor,
                 int32_t n_iter, int32_t assemble) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!f || !f->P || n_iter < 0) return fail(TSE_ERR_ARG, "invalid argument");
    if (assemble) {
        rc = launch_p_assemble(ctx, f, dt);
        if (rc) return rc;
    } else if (ctx->rec_elems < rec_buffer_elems(ctx->geo, kPRec) ||
               ctx->mask_elems < mask_bytes(ctx->geo)) {
        return fail(TSE_ERR_STATE, "tse_p_sweeps: no records assembled");
    }
    for (int n = 0; n < n_iter; ++n)
        for (int colour = 0; colour < 2; ++colour) {
            rc = launch_p_sweep(ctx, f->P, colour, sor,
                                SweepCtl{nullptr, nullptr, nullptr, ctx->partials, 0, 0});
            if (rc) return rc;
        }
    return 0;
}

}  // extern "C"
