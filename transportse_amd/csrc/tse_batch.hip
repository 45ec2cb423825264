// tse_batch.hip -- batched performance-form solves over many contexts.
//
// A batch holds member slots; each member is a context whose fast system of
// the batch's kind (pressure or transport) is set up.  One cycle runs, for
// every active member, exactly what the single-system solve runs per cycle:
// `inner` fused sweeps (FP64 on the iterate, or FP32 on the correction from
// zero), then the FP64 residual pass, then the convergence test against the
// member's own tolerance.  Members therefore follow the same arithmetic
// whether they are solved alone or together with others.
//
// Sweeps: one persistent launch per NS-sweep group over the tiles of all
// members.  Work items (member, tile) sit in a device list; each block takes
// the next item from an atomic counter until the list is exhausted.  Every
// block makes exactly one fetch past the end, so launch l of a cycle starts
// its counter at l * (items + blocks) after the cycle's reset kernel; the
// counter never needs a host-side memset and the cycle can be captured as a
// hipGraph.  Per-member state (ping-pong index, active flag, sweep count,
// residual) lives in device memory and is updated by the cycle's reduce
// kernel, so a captured cycle stays valid while the member set is unchanged.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tse_fast.h"

namespace tse {

constexpr int kChunk = kBlock * 16;  // cells per residual work item

struct BMember {
    ZGeo g;
    FastDiv dnx, dny;
    int nx, ny, nz, rchunk0;
    long N;
    double sor;     // this member's over-relaxation factor
    double *p[2];   // FP64 iterate ping-pong
    float *e[2];    // FP32 correction ping-pong
    float *r32;
    FastOp fop;
    TrOp top;
};

struct BStat {      // per-slot state, device-resident, mirrored to the host
    int active;     // takes part in the next cycle
    int cur;        // which p buffer holds the iterate
    int sweeps;
    int cycles;
    double bnorm;   // ||b|| (1 if b = 0)
    double rel;     // last relative residual
    double rel0;    // initial relative residual
    double rel_prev;  // relative residual before the last cycle (0: none)
};

__global__ void batch_reset_kernel(unsigned *queue) {
    if (threadIdx.x == 0) queue[0] = 0u;
}

__device__ inline int fetch_item(unsigned *queue, unsigned base) {
    __shared__ int s_item;
    if (threadIdx.x == 0 && threadIdx.y == 0)
        s_item = (int)(atomicAdd(queue, 1u) - base);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    return it;
}

// Persistent pressure z-march over (member, tile) items.  correction != 0:
// FP32 sweeps on e (launch 0 starts from zero), else FP64 sweeps on p.
template <typename T, int NS, int BYT, int RY, int VAR>
__global__ void __launch_bounds__(64 * BYT)
batch_rb_zmarch_kernel(const BMember *__restrict__ mem,
                       const BStat *__restrict__ st,
                       const int2 *__restrict__ items, int n_items,
                       unsigned *queue, unsigned base, int launch) {
    for (;;) {
        const int it = fetch_item(queue, base);
        if (it >= n_items) break;
        const int2 w = items[it];
        const BMember &m = mem[w.x];
        if (!st[w.x].active) continue;  // uniform per block
        const T omega = (T)m.sor;
        const T *pin;
        T *pout;
        const T *wx, *wy, *wz, *inv, *rhs;
        int zero;
        if constexpr (sizeof(T) == 4) {
            pin = m.e[launch & 1];
            pout = m.e[(launch + 1) & 1];
            wx = m.fop.Wx32;
            wy = m.fop.Wy32;
            wz = m.fop.Wz32;
            inv = m.fop.invD32;
            rhs = m.r32;
            zero = launch == 0;
        } else {
            const int c = (st[w.x].cur + launch) & 1;
            pin = m.p[c];
            pout = m.p[c ^ 1];
            wx = m.fop.Wx;
            wy = m.fop.Wy;
            wz = m.fop.Wz;
            inv = m.fop.invD;
            rhs = m.fop.b;
            zero = 0;
        }
        if constexpr (VAR == 0)
            rb_zmarch_tile<T, NS, BYT, RY>(m.g, w.y, pin, pout, wx, wy, wz, inv,
                                           rhs, omega, zero);
        else if constexpr (VAR == 1)
            rb_zmarch_tile_xl<T, NS, BYT, RY>(m.g, w.y, pin, pout, wx, wy, wz, inv,
                                              rhs, omega, zero);
        else if constexpr (VAR == 2)
            rb_zmarch_tile_xr<T, NS, BYT, RY>(m.g, w.y, pin, pout, wx, wy, wz, inv,
                                              rhs, omega, zero);
        else if constexpr (VAR == 3) {
            static_assert(sizeof(T) == 4 && RY == 2, "variant 3: FP32, cell pairs");
            rb_zmarch_tile_x2<NS, BYT, false>(m.g, w.y, pin, pout, wx, wy, wz, inv, rhs,
                                              omega, zero);
        } else {
            // the cell-pair sweep on the bfloat16 inner operator
            static_assert(sizeof(T) == 4 && RY == 2 && VAR == 5, "variant 5: FP32, cell pairs");
            rb_zmarch_tile_x5<NS, BYT, false>(m.g, w.y, pin, pout,
                                       (const uint32_t *)m.fop.rec16, rhs, omega, zero);
        }
        __syncthreads();  // LDS reuse by the next tile
    }
}

template <typename T, int NS, int BY, int VAR>
__global__ void __launch_bounds__(64 * BY)
batch_tr_zmarch_kernel(const BMember *__restrict__ mem,
                       const BStat *__restrict__ st,
                       const int2 *__restrict__ items, int n_items,
                       unsigned *queue, unsigned base, int launch) {
    for (;;) {
        const int it = fetch_item(queue, base);
        if (it >= n_items) break;
        const int2 w = items[it];
        const BMember &m = mem[w.x];
        if (!st[w.x].active) continue;
        const T omega = (T)m.sor;
        const T *pin, *coef, *inv, *rhs;
        T *pout;
        int zero;
        if constexpr (sizeof(T) == 4) {
            pin = m.e[launch & 1];
            pout = m.e[(launch + 1) & 1];
            coef = m.top.c32;
            inv = m.top.invD32;
            rhs = m.r32;
            zero = launch == 0;
        } else {
            const int c = (st[w.x].cur + launch) & 1;
            pin = m.p[c];
            pout = m.p[c ^ 1];
            coef = m.top.c;
            inv = m.top.invD;
            rhs = m.top.b;
            zero = 0;
        }
        if constexpr (VAR == 0)
            tr_zmarch_tile<T, NS, BY>(m.g, w.y, pin, pout, coef, m.top.ld, inv, rhs,
                                      omega, zero);
        else
            tr_zmarch_tile_xr<T, NS, BY>(m.g, w.y, pin, pout, coef, m.top.ld, inv,
                                         rhs, omega, zero);
        __syncthreads();
    }
}

// Residual over (member, chunk) items, one item per block; block b handles
// item first + b.  MODE as p_residual_cell.  The iterate read is
// p[cur ^ pshift]; mode 1 writes p[cur ^ 1] = p[cur] + e[ecur].  `init`
// selects members regardless of their active flag (first residual).
template <int KIND, int MODE>
__global__ void __launch_bounds__(kBlock)
batch_residual_kernel(const BMember *__restrict__ mem,
                      const BStat *__restrict__ st,
                      const int2 *__restrict__ ritems, int first, int pshift,
                      int ecur, int init, double *__restrict__ partials) {
    const int it = first + blockIdx.x;
    const int2 w = ritems[it];
    if (!init && !st[w.x].active) return;
    const BMember &m = mem[w.x];
    const int c = st[w.x].cur ^ pshift;
    const double *pin = m.p[c];
    double *pout = m.p[c ^ 1];
    const float *e = m.e[ecur];
    const long t0 = (long)w.y * kChunk;
    const long t1 = std::min(t0 + kChunk, m.N);
    double sr = 0.0, sb = 0.0;
    for (long t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
        const int row = (int)fdiv((uint32_t)t, m.dnx);
        const int i = (int)t - row * m.nx;
        const int k = (int)fdiv((uint32_t)row, m.dny);
        const int j = row - k * m.ny;
        if (KIND == SYS_P)
            p_residual_cell<MODE>(m.nx, m.ny, m.nz, i, j, k, t, pin, e, pout,
                                  m.fop, m.r32, sr, sb);
        else
            tr_residual_cell<MODE>(m.nx, m.ny, m.nz, i, j, k, t, pin, e, pout,
                                   m.top, m.r32, sr, sb);
    }
    block_sum2_to(sr, sb, partials + 2 * it);
}

// One block per slot (slot = first + blockIdx.x): sums the member's residual
// partials and updates its state.  init: first residual of a new member.
__global__ void __launch_bounds__(kBlock)
batch_reduce_kernel(const BMember *__restrict__ mem, BStat *__restrict__ st,
                    const unsigned char *__restrict__ used, int first,
                    const double *__restrict__ partials, int init, int flip,
                    int inner, int max_sweeps, double tol) {
    const int s = first + blockIdx.x;
    if (!used[s]) return;
    if (!init && !st[s].active) return;
    const BMember &m = mem[s];
    const int nch = (int)((m.N + kChunk - 1) / kChunk);
    double a = 0, c = 0;
    for (int q = threadIdx.x; q < nch; q += blockDim.x) {
        a += partials[2 * (m.rchunk0 + q)];
        c += partials[2 * (m.rchunk0 + q) + 1];
    }
    __shared__ double out[2];
    block_sum2_to(a, c, out);
    __syncthreads();
    if (threadIdx.x == 0) {
        BStat x = st[s];
        const double rn = sqrt(out[0]);
        if (init) {
            const double bn = sqrt(out[1]);
            x.bnorm = bn > 0 ? bn : 1.0;
            x.rel = rn / x.bnorm;
            x.rel0 = x.rel;
            x.rel_prev = 0.0;
            x.sweeps = 0;
            x.cycles = 0;
        } else {
            x.rel_prev = x.rel;
            x.rel = rn / x.bnorm;
            x.sweeps += inner;
            x.cycles += 1;
            x.cur ^= flip;
        }
        x.active = (x.rel > tol && x.sweeps < max_sweeps && x.rel == x.rel) ? 1 : 0;
        st[s] = x;
    }
}

}  // namespace tse

using namespace tse;

struct tse_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int kind = SYS_P, form = TSE_FORM_MIXED;
    tse_sweep_plan plan{2, 16, 1, 0};
    int inner = 100, max_sweeps = 100000;
    double sor = 1.9, tol = 1e-10;
    bool begun = false;
    int cap = 0;                      // slot capacity
    std::vector<tse_ctx *> ctx;       // per slot; nullptr = free
    std::vector<BMember> hmem;
    std::vector<unsigned char> hused;
    BMember *dmem = nullptr;
    BStat *dst = nullptr;
    unsigned char *dused = nullptr;
    BStat *hst = nullptr;             // pinned mirror
    int2 *ditems = nullptr;
    int items_cap = 0, n_items = 0;
    int2 *dritems = nullptr;
    int ritems_cap = 0, n_ritems = 0;
    double *dpartials = nullptr;
    unsigned *dqueue = nullptr;
    int grid = 0;                     // persistent blocks per sweep launch
    bool dirty = true;                // items / graph need a rebuild
    bool sor_dirty = false;           // hmem's sor values not yet on the device
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int batch_fail_hip(hipError_t e, const char *what) {
    return fail(TSE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define TSE_B_TRY(expr)                                          \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return batch_fail_hip(e_, #expr);  \
    } while (0)

void drop_graph(tse_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int grow_slots(tse_batch *b, int need) {
    if (need <= b->cap) return 0;
    int cap = std::max(16, b->cap);
    while (cap < need) cap *= 2;
    BMember *dm;
    BStat *ds, *hs;
    unsigned char *du;
    TSE_B_TRY(hipMalloc((void **)&dm, sizeof(BMember) * cap));
    TSE_B_TRY(hipMalloc((void **)&ds, sizeof(BStat) * cap));
    TSE_B_TRY(hipMalloc((void **)&du, cap));
    TSE_B_TRY(hipHostMalloc((void **)&hs, sizeof(BStat) * cap));
    TSE_B_TRY(hipStreamSynchronize(b->stream));
    if (b->cap) {
        TSE_B_TRY(hipMemcpy(ds, b->dst, sizeof(BStat) * b->cap, hipMemcpyDeviceToDevice));
        std::copy(b->hst, b->hst + b->cap, hs);
        (void)hipFree(b->dmem);
        (void)hipFree(b->dst);
        (void)hipFree(b->dused);
        (void)hipHostFree(b->hst);
    }
    b->dmem = dm;
    b->dst = ds;
    b->dused = du;
    b->hst = hs;
    b->ctx.resize(cap, nullptr);
    b->hmem.resize(cap);
    b->hused.resize(cap, 0);
    b->cap = cap;
    b->dirty = true;
    return 0;
}

ZGeo member_zgeo(const tse_batch *b, const tse_ctx *c) {
    return make_zgeo(fast_tiling_geo(c->geo, b->kind), b->plan.ns, b->plan.by,
                     b->plan.zchunk, b->plan.variant >= 2);
}

// Rebuild work-item lists, member table and the persistent grid size.
int rebuild(tse_batch *b) {
    std::vector<int2> items, ritems;
    for (int s = 0; s < b->cap; ++s) {
        if (!b->hused[s]) continue;
        BMember &m = b->hmem[s];
        m.rchunk0 = (int)ritems.size();
        const int nch = (int)((m.N + kChunk - 1) / kChunk);
        for (int q = 0; q < nch; ++q) ritems.push_back(make_int2(s, q));
        for (int t = 0; t < m.g.ntiles; ++t) items.push_back(make_int2(s, t));
    }
    // Longest tiles first (plane steps of the z-march: the chunk plus its
    // halo planes), so a launch does not end on a few tall columns; tiles
    // write disjoint cells from the launch's input buffer, so the order
    // does not change any result.
    const int R = 2 * b->plan.ns;
    auto steps = [&](const int2 &w) {
        const ZGeo &g = b->hmem[w.x].g;
        const int k0 = (w.y / (g.tiles_x * g.tiles_y)) * g.zchunk;
        const int k1 = std::min(k0 + g.zchunk, g.nz);
        return (k1 + R) - std::max(k0 - R, 0);
    };
    std::stable_sort(items.begin(), items.end(),
                     [&](const int2 &a, const int2 &c) { return steps(a) > steps(c); });
    if ((int)items.size() > b->items_cap) {
        if (b->ditems) (void)hipFree(b->ditems);
        b->items_cap = std::max<int>(items.size(), 2 * b->items_cap);
        TSE_B_TRY(hipMalloc((void **)&b->ditems, sizeof(int2) * b->items_cap));
    }
    if ((int)ritems.size() > b->ritems_cap) {
        if (b->dritems) (void)hipFree(b->dritems);
        if (b->dpartials) (void)hipFree(b->dpartials);
        b->ritems_cap = std::max<int>(ritems.size(), 2 * b->ritems_cap);
        TSE_B_TRY(hipMalloc((void **)&b->dritems, sizeof(int2) * b->ritems_cap));
        TSE_B_TRY(hipMalloc((void **)&b->dpartials, sizeof(double) * 2 * b->ritems_cap));
    }
    b->n_items = (int)items.size();
    b->n_ritems = (int)ritems.size();
    if (b->n_items)
        TSE_B_TRY(hipMemcpyAsync(b->ditems, items.data(), sizeof(int2) * items.size(),
                                 hipMemcpyHostToDevice, b->stream));
    if (b->n_ritems)
        TSE_B_TRY(hipMemcpyAsync(b->dritems, ritems.data(), sizeof(int2) * ritems.size(),
                                 hipMemcpyHostToDevice, b->stream));
    TSE_B_TRY(hipMemcpyAsync(b->dmem, b->hmem.data(), sizeof(BMember) * b->cap,
                             hipMemcpyHostToDevice, b->stream));
    TSE_B_TRY(hipMemcpyAsync(b->dused, b->hused.data(), b->cap,
                             hipMemcpyHostToDevice, b->stream));
    TSE_B_TRY(hipStreamSynchronize(b->stream));  // host vectors go out of scope
    drop_graph(b);
    b->dirty = false;
    return 0;
}

template <typename K>
int occupancy_grid(K kernel, int threads, int n_items, int *grid) {
    int dev, per_cu = 0;
    TSE_B_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    TSE_B_TRY(hipGetDeviceProperties(&prop, dev));
    TSE_B_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    *grid = std::max(1, std::min(n_items, std::max(1, per_cu) * prop.multiProcessorCount));
    return 0;
}

// Launch one persistent sweep launch (index `launch` within the cycle).
template <typename T>
int launch_sweeps(tse_batch *b, int launch, unsigned base, bool size_only) {
    const int ns = b->plan.ns, by = b->plan.by, ry = b->plan.ry, v = b->plan.variant;
#define TSE_BP(V_, NS_, BYT_, RY_)                                                     \
    if (b->kind == SYS_P && v == V_ && ns == NS_ && by == BYT_ * RY_ && ry == RY_) {   \
        auto k = batch_rb_zmarch_kernel<T, NS_, BYT_, RY_, V_>;                        \
        if (size_only) return occupancy_grid(k, 64 * BYT_, b->n_items, &b->grid);     \
        hipLaunchKernelGGL(k, dim3(b->grid), dim3(64, BYT_), 0, b->stream, b->dmem,    \
                           b->dst, b->ditems, b->n_items, b->dqueue, base, launch);    \
        TSE_B_TRY(hipGetLastError());                                                  \
        return 0;                                                                      \
    }
#define TSE_BT(V_, NS_, BY_)                                                           \
    if (b->kind == SYS_TR && v == V_ && ns == NS_ && by == BY_ && ry == 1) {           \
        auto k = batch_tr_zmarch_kernel<T, NS_, BY_, V_>;                              \
        if (size_only) return occupancy_grid(k, 64 * BY_, b->n_items, &b->grid);      \
        hipLaunchKernelGGL(k, dim3(b->grid), dim3(64, BY_), 0, b->stream, b->dmem,     \
                           b->dst, b->ditems, b->n_items, b->dqueue, base, launch);    \
        TSE_B_TRY(hipGetLastError());                                                  \
        return 0;                                                                      \
    }
    TSE_BP(0, 1, 8, 1)
    TSE_BP(0, 1, 16, 1)
    TSE_BP(0, 2, 16, 1)
    TSE_BP(0, 1, 12, 2)
    TSE_BP(1, 1, 16, 1)
    TSE_BP(1, 2, 16, 1)
    TSE_BP(1, 1, 12, 2)
    TSE_BP(2, 1, 16, 1)
    TSE_BP(2, 2, 16, 1)
    TSE_BP(2, 1, 12, 2)
    TSE_BT(0, 1, 8)
    TSE_BT(0, 1, 16)
    TSE_BT(0, 2, 16)
    TSE_BT(2, 1, 8)
    TSE_BT(2, 1, 16)
    TSE_BT(2, 2, 16)
    if constexpr (sizeof(T) == 4) {
        TSE_BP(0, 1, 16, 2)
        TSE_BP(0, 2, 16, 2)
        TSE_BP(0, 2, 8, 4)
        TSE_BP(0, 3, 16, 1)
        TSE_BP(1, 2, 12, 2)
        TSE_BP(1, 2, 16, 2)
        TSE_BP(1, 3, 16, 1)
        TSE_BP(1, 3, 12, 2)
        TSE_BP(2, 2, 12, 2)
        TSE_BP(2, 2, 16, 2)
        TSE_BP(2, 3, 16, 1)
        TSE_BP(2, 3, 12, 2)
        TSE_BP(2, 3, 16, 2)
        TSE_BP(3, 2, 16, 2)
        TSE_BP(3, 3, 16, 2)
        TSE_BP(5, 2, 16, 2)
        TSE_BP(5, 3, 16, 2)
        TSE_BT(0, 3, 16)
        TSE_BT(2, 3, 16)
    }
#undef TSE_BP
#undef TSE_BT
    return fail(TSE_ERR_ARG, "unsupported batch sweep plan");
}

template <int KIND>
int launch_residual(tse_batch *b, int mode, int first, int count, int pshift,
                    int ecur, int init) {
    if (count <= 0) return 0;
#define TSE_BR(MODE_)                                                                 \
    if (mode == MODE_) {                                                              \
        hipLaunchKernelGGL((batch_residual_kernel<KIND, MODE_>), dim3(count),         \
                           dim3(kBlock), 0, b->stream, b->dmem, b->dst, b->dritems,   \
                           first, pshift, ecur, init, b->dpartials);                  \
        TSE_B_TRY(hipGetLastError());                                                 \
        return 0;                                                                     \
    }
    TSE_BR(0)
    TSE_BR(1)
    TSE_BR(2)
#undef TSE_BR
    return fail(TSE_ERR_ARG, "bad residual mode");
}

int residual_any(tse_batch *b, int mode, int first, int count, int pshift,
                 int ecur, int init) {
    return b->kind == SYS_P
               ? launch_residual<SYS_P>(b, mode, first, count, pshift, ecur, init)
               : launch_residual<SYS_TR>(b, mode, first, count, pshift, ecur, init);
}

// Enqueue one cycle for all active members (capturable: no host reads).
int enqueue_cycle(tse_batch *b) {
    const bool fp32 = b->form == TSE_FORM_MIXED;
    const int L = b->inner / b->plan.ns;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, b->stream, b->dqueue);
    TSE_B_TRY(hipGetLastError());
    for (int l = 0; l < L; ++l) {
        const unsigned base = (unsigned)l * (unsigned)(b->n_items + b->grid);
        const int rc = fp32 ? launch_sweeps<float>(b, l, base, false)
                            : launch_sweeps<double>(b, l, base, false);
        if (rc) return rc;
    }
    int rc = fp32 ? residual_any(b, 1, 0, b->n_ritems, 0, L & 1, 0)
                  : residual_any(b, 0, 0, b->n_ritems, L & 1, 0, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(b->cap), dim3(kBlock), 0, b->stream,
                       b->dmem, b->dst, b->dused, 0, b->dpartials, 0,
                       fp32 ? 1 : (L & 1), b->inner, b->max_sweeps, b->tol);
    TSE_B_TRY(hipGetLastError());
    TSE_B_TRY(hipMemcpyAsync(b->hst, b->dst, sizeof(BStat) * b->cap,
                             hipMemcpyDeviceToHost, b->stream));
    return 0;
}

int prepare(tse_batch *b) {
    if (!b->dirty) {
        if (b->sor_dirty) {   // member table from host memory, before the next sweeps
            TSE_B_TRY(hipMemcpyAsync(b->dmem, b->hmem.data(), sizeof(BMember) * b->cap,
                                     hipMemcpyHostToDevice, b->stream));
            b->sor_dirty = false;
        }
        return 0;
    }
    int rc = rebuild(b);   // uploads the whole member table
    if (rc) return rc;
    b->sor_dirty = false;
    if (b->n_items == 0) return 0;
    const bool fp32 = b->form == TSE_FORM_MIXED;
    return fp32 ? launch_sweeps<float>(b, 0, 0, true)
                : launch_sweeps<double>(b, 0, 0, true);
}

int check_batch(tse_batch *b) {
    if (!b) return fail(TSE_ERR_ARG, "null batch");
    if (!b->begun) return fail(TSE_ERR_STATE, "tse_batch_begin not called");
    return hipSetDevice(b->device) == hipSuccess ? 0 : fail(TSE_ERR_HIP, "hipSetDevice");
}

// Member table entry for context c's fast system s.
BMember member_entry(const tse_batch *b, tse_ctx *c, FastSys *s) {
    const Geo &G = c->geo;
    BMember m{};
    m.g = member_zgeo(b, c);
    m.dnx = make_fastdiv(s->nxf);
    m.dny = make_fastdiv(G.ny);
    m.nx = s->nxf;   // row length of the natural layout
    m.ny = G.ny;
    m.nz = G.nz;
    m.N = s->N;
    m.p[0] = s->p(0);
    m.p[1] = s->p(1);
    m.e[0] = s->e(0);
    m.e[1] = s->e(1);
    m.r32 = s->r32();
    if (b->kind == SYS_P) m.fop = s->pop();
    else m.top = s->trop();
    m.sor = b->sor;
    return m;
}

FastSys *member_sys(const tse_batch *b, tse_ctx *c) {
    return (FastSys *)(b->kind == SYS_P ? c->fast : c->fast_tr);
}

// (Re)starts the solve of the member in `slot` from its context's iterate:
// uploads its ping-pong index and evaluates its initial residual.
int init_member(tse_batch *b, int slot) {
    tse_ctx *c = b->ctx[slot];
    FastSys *s = member_sys(b, c);
    if (c->stream != b->stream) TSE_B_TRY(hipStreamSynchronize(c->stream));
    BMember m = member_entry(b, c, s);
    const BMember &o = b->hmem[slot];
    m.sor = o.sor;
    if (m.N != o.N || m.p[0] != o.p[0] || m.p[1] != o.p[1] || m.e[0] != o.e[0] ||
        m.nx != o.nx || m.ny != o.ny) {
        b->hmem[slot] = m;
        b->dirty = true;
    }
    BStat st0{};
    st0.cur = s->cur;
    b->hst[slot] = st0;
    TSE_B_TRY(hipMemcpyAsync(b->dst + slot, &b->hst[slot], sizeof(BStat),
                             hipMemcpyHostToDevice, b->stream));
    int rc = prepare(b);
    if (rc) return rc;
    const BMember &mm = b->hmem[slot];
    const int nch = (int)((mm.N + kChunk - 1) / kChunk);
    const bool fp32 = b->form == TSE_FORM_MIXED;
    rc = residual_any(b, fp32 ? 2 : 0, mm.rchunk0, nch, 0, 0, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(1), dim3(kBlock), 0, b->stream,
                       b->dmem, b->dst, b->dused, slot, b->dpartials, 1, 0,
                       b->inner, b->max_sweeps, b->tol);
    TSE_B_TRY(hipGetLastError());
    TSE_B_TRY(hipMemcpyAsync(&b->hst[slot], b->dst + slot, sizeof(BStat),
                             hipMemcpyDeviceToHost, b->stream));
    TSE_B_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

}  // namespace

extern "C" {

int tse_batch_create(tse_batch **out, int32_t device, void *stream) {
    if (!out) return fail(TSE_ERR_ARG, "null out");
    if (hipSetDevice(device) != hipSuccess) return fail(TSE_ERR_HIP, "hipSetDevice");
    tse_batch *b = new tse_batch();
    b->device = device;
    b->stream = (hipStream_t)stream;
    if (hipMalloc((void **)&b->dqueue, sizeof(unsigned)) != hipSuccess) {
        delete b;
        return fail(TSE_ERR_NOMEM, "batch queue allocation failed");
    }
    *out = b;
    return 0;
}

int tse_batch_destroy(tse_batch *b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->dmem) (void)hipFree(b->dmem);
    if (b->dst) (void)hipFree(b->dst);
    if (b->dused) (void)hipFree(b->dused);
    if (b->hst) (void)hipHostFree(b->hst);
    if (b->ditems) (void)hipFree(b->ditems);
    if (b->dritems) (void)hipFree(b->dritems);
    if (b->dpartials) (void)hipFree(b->dpartials);
    if (b->dqueue) (void)hipFree(b->dqueue);
    delete b;
    return 0;
}

int tse_batch_begin(tse_batch *b, int32_t kind, const tse_fast_params *prm) {
    if (!b || !prm) return fail(TSE_ERR_ARG, "null argument");
    if (kind != SYS_P && kind != SYS_TR) return fail(TSE_ERR_ARG, "kind must be 0 (P) or 1 (TR)");
    const int fp32 = prm->form == TSE_FORM_MIXED;
    if (prm->form != TSE_FORM_MIXED && prm->form != TSE_FORM_PLAIN64)
        return fail(TSE_ERR_ARG, "form must be PLAIN64 or MIXED");
    const tse_sweep_plan &pl = prm->plan;
    if (!fast_plan_ok(pl, fp32, kind) || prm->inner < pl.ns ||
        prm->inner % pl.ns || prm->max_sweeps < 0)
        return fail(TSE_ERR_ARG, "unsupported plan / inner sweep count");
    if (pl.variant == 7)   // the batch residual writes an FP32 rhs
        return fail(TSE_ERR_ARG, "sweep variant 7 runs single systems only");
    if (prm->adapt_sor)    // members keep their own factor (tse_batch_set_sor)
        return fail(TSE_ERR_ARG, "adapt_sor is for single solves");
    for (int s = 0; s < b->cap; ++s)
        if (b->hused[s]) return fail(TSE_ERR_STATE, "batch still has members");
    b->kind = kind;
    b->form = prm->form;
    b->plan = prm->plan;
    b->inner = prm->inner;
    b->max_sweeps = prm->max_sweeps;
    b->sor = prm->sor;
    b->tol = prm->tol;
    b->begun = true;
    b->dirty = true;
    return 0;
}

// Adds a member (its fast system of the batch's kind must be set up) and
// evaluates its initial residual.  Returns the slot (>= 0) or an error (< 0).
int tse_batch_add(tse_batch *b, tse_ctx *c) {
    int rc = check_batch(b);
    if (rc) return rc;
    if (!c || !c->grid_set) return fail(TSE_ERR_ARG, "member context without grid");
    if (c->device != b->device) return fail(TSE_ERR_ARG, "member on another device");
    FastSys *s = member_sys(b, c);
    if (!s || !s->ready) return fail(TSE_ERR_STATE, "member's fast system not set up");
    if (!fast_plan_fits(b->plan, s->N, b->form == TSE_FORM_MIXED))
        return fail(TSE_ERR_ARG, "sweep variants 2, 3 and 5 need fields of at most 1 GiB "
                                 "(variant 5: 8 bytes per cell)");
    if (!fast_plan_grid_ok(b->plan, s->nxf))
        return fail(TSE_ERR_ARG, "sweep variants 3 and 5 need an even nx");
    for (int q = 0; q < b->cap; ++q)
        if (b->hused[q] && b->ctx[q] == c)
            return fail(TSE_ERR_ARG, "context already in the batch");
    int slot = -1;
    for (int q = 0; q < b->cap; ++q)
        if (!b->hused[q]) { slot = q; break; }
    if (slot < 0) {
        slot = b->cap;
        rc = grow_slots(b, b->cap + 1);
        if (rc) return rc;
    }
    b->ctx[slot] = c;
    b->hmem[slot] = member_entry(b, c, s);
    b->hused[slot] = 1;
    b->dirty = true;
    rc = init_member(b, slot);
    return rc ? rc : slot;
}

// Restarts a member after its fast system was set up again (new operator
// and iterate, e.g. the next time step).  The captured cycle stays valid
// while the member's grid and buffers are unchanged.
int tse_batch_rearm(tse_batch *b, int32_t slot) {
    int rc = check_batch(b);
    if (rc) return rc;
    if (slot < 0 || slot >= b->cap || !b->hused[slot])
        return fail(TSE_ERR_ARG, "no member in that slot");
    FastSys *s = member_sys(b, b->ctx[slot]);
    if (!s || !s->ready) return fail(TSE_ERR_STATE, "member's fast system not set up");
    return init_member(b, slot);
}

// Re-arms several members at once: one synchronisation for all of them
// (each member's initial residual still runs over its own chunks).
int tse_batch_rearm_many(tse_batch *b, const int32_t *slots, int32_t n) {
    int rc = check_batch(b);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !slots)) return fail(TSE_ERR_ARG, "bad slot list");
    for (int q = 0; q < n; ++q) {
        const int slot = slots[q];
        if (slot < 0 || slot >= b->cap || !b->hused[slot])
            return fail(TSE_ERR_ARG, "no member in that slot");
        tse_ctx *c = b->ctx[slot];
        FastSys *s = member_sys(b, c);
        if (!s || !s->ready) return fail(TSE_ERR_STATE, "member's fast system not set up");
        if (c->stream != b->stream) TSE_B_TRY(hipStreamSynchronize(c->stream));
        BMember m = member_entry(b, c, s);
        const BMember &o = b->hmem[slot];
        m.sor = o.sor;
        if (m.N != o.N || m.p[0] != o.p[0] || m.p[1] != o.p[1] ||
            m.e[0] != o.e[0] || m.nx != o.nx || m.ny != o.ny) {
            b->hmem[slot] = m;
            b->dirty = true;
        }
    }
    rc = prepare(b);
    if (rc) return rc;
    const bool fp32 = b->form == TSE_FORM_MIXED;
    // hst mirrors the device state of every slot (each cycle and re-arm ends
    // with a read-back, and nothing runs on the batch between calls), so the
    // re-armed entries go up with the others in one copy
    for (int q = 0; q < n; ++q) {
        const int slot = slots[q];
        BStat st0{};
        st0.cur = member_sys(b, b->ctx[slot])->cur;
        b->hst[slot] = st0;
    }
    if (n > 0)
        TSE_B_TRY(hipMemcpyAsync(b->dst, b->hst, sizeof(BStat) * b->cap,
                                 hipMemcpyHostToDevice, b->stream));
    // every member re-armed (the lockstep case): one initial-residual launch
    // over all items and one reduction over all slots; per-member sums are
    // the same as with a launch per member
    std::vector<char> mark(b->cap, 0);
    int distinct = 0, used = 0;
    for (int q = 0; q < n; ++q)
        if (!mark[slots[q]]) {
            mark[slots[q]] = 1;
            ++distinct;
        }
    for (int q = 0; q < b->cap; ++q) used += b->hused[q] ? 1 : 0;
    if (n > 0 && distinct == used && b->n_ritems > 0) {
        rc = residual_any(b, fp32 ? 2 : 0, 0, b->n_ritems, 0, 0, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(batch_reduce_kernel, dim3(b->cap), dim3(kBlock), 0, b->stream,
                           b->dmem, b->dst, b->dused, 0, b->dpartials, 1, 0,
                           b->inner, b->max_sweeps, b->tol);
        TSE_B_TRY(hipGetLastError());
        n = 0;   // done
    }
    for (int q = 0; q < n; ++q) {
        const int slot = slots[q];
        const BMember &mm = b->hmem[slot];
        const int nch = (int)((mm.N + kChunk - 1) / kChunk);
        rc = residual_any(b, fp32 ? 2 : 0, mm.rchunk0, nch, 0, 0, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(batch_reduce_kernel, dim3(1), dim3(kBlock), 0, b->stream,
                           b->dmem, b->dst, b->dused, slot, b->dpartials, 1, 0,
                           b->inner, b->max_sweeps, b->tol);
        TSE_B_TRY(hipGetLastError());
    }
    TSE_B_TRY(hipMemcpyAsync(b->hst, b->dst, sizeof(BStat) * b->cap,
                             hipMemcpyDeviceToHost, b->stream));
    TSE_B_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

// Hands the member's current iterate back to its context (the fast-form
// result / residual calls then see it) without removing it.  Host only: every
// batch call that launches work (add, re-arm, cycle, run) ends by reading the
// state back after its stream has drained, so hst is current and the
// batch's kernels are done; work the caller has queued since (a member's next
// set-up on the same stream) is not waited for.
int tse_batch_writeback(tse_batch *b, int32_t slot) {
    int rc = check_batch(b);
    if (rc) return rc;
    if (slot < 0 || slot >= b->cap || !b->hused[slot])
        return fail(TSE_ERR_ARG, "no member in that slot");
    member_sys(b, b->ctx[slot])->cur = b->hst[slot].cur;
    return 0;
}

// Sets the over-relaxation factor of one member (default: the batch's).
// Takes effect from the next cycle or re-arm, which uploads the member table
// once for all changed members; a captured cycle stays valid.
int tse_batch_set_sor(tse_batch *b, int32_t slot, double sor) {
    int rc = check_batch(b);
    if (rc) return rc;
    if (slot < 0 || slot >= b->cap || !b->hused[slot])
        return fail(TSE_ERR_ARG, "no member in that slot");
    if (!(sor > 0.0 && sor < 2.0)) return fail(TSE_ERR_ARG, "sor must be in (0, 2)");
    b->hmem[slot].sor = sor;
    b->sor_dirty = true;
    return 0;
}

// Removes a member and hands its iterate back to its context.
int tse_batch_remove(tse_batch *b, int32_t slot) {
    int rc = tse_batch_writeback(b, slot);
    if (rc) return rc;
    b->ctx[slot] = nullptr;
    b->hused[slot] = 0;
    b->dirty = true;
    return 0;
}

// One cycle for every active member; *n_active receives the number of
// members still active afterwards.  use_graph != 0 captures the cycle as a
// hipGraph on first use and replays it while the member set is unchanged.
int tse_batch_cycle(tse_batch *b, int32_t use_graph, int32_t *n_active) {
    int rc = check_batch(b);
    if (rc) return rc;
    rc = prepare(b);
    if (rc) return rc;
    int act = 0;
    for (int s = 0; s < b->cap; ++s) act += b->hused[s] && b->hst[s].active;
    if (act && b->n_items) {
        if (use_graph) {
            if (!b->exec) {
                TSE_B_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                rc = enqueue_cycle(b);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(b->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                TSE_B_TRY(e);
                b->graph = g;
                TSE_B_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            }
            TSE_B_TRY(hipGraphLaunch(b->exec, b->stream));
        } else {
            rc = enqueue_cycle(b);
            if (rc) return rc;
        }
        TSE_B_TRY(hipStreamSynchronize(b->stream));
        act = 0;
        for (int s = 0; s < b->cap; ++s) act += b->hused[s] && b->hst[s].active;
    }
    if (n_active) *n_active = act;
    return 0;
}

// Cycles until no member is active; *cycles receives the number run.
int tse_batch_run(tse_batch *b, int32_t use_graph, int32_t *cycles) {
    int n = 0, act = 1, rc;
    while (act > 0) {
        rc = tse_batch_cycle(b, use_graph, &act);
        if (rc) return rc;
        ++n;
    }
    if (cycles) *cycles = n;
    return 0;
}

int tse_batch_status(tse_batch *b, int32_t slot, tse_fast_result *out) {
    int rc = check_batch(b);
    if (rc) return rc;
    if (!out || slot < 0 || slot >= b->cap || !b->hused[slot])
        return fail(TSE_ERR_ARG, "no member in that slot");
    const BStat &x = b->hst[slot];
    out->sweeps = x.sweeps;
    out->cycles = x.cycles;
    out->converged = x.rel <= b->tol ? 1 : 0;
    out->pad = x.active;
    out->initial_rel = x.rel0;
    out->rel_residual = x.rel;
    // per sweep over the last cycle after the first (every cycle of a batch
    // runs `inner` sweeps), as a single solve reports it: only from a cycle
    // that spans inner |log(sor - 1)| >= 1/4
    const double sor = b->hmem[slot].sor;
    const bool resolved = sor > 1.0 && sor < 2.0 &&
                          b->inner * std::fabs(std::log(sor - 1.0)) >= 0.25;
    out->contraction =
        resolved && x.cycles > 1 && x.rel_prev > 0 && x.rel > 0 && x.rel < x.rel_prev
            ? std::exp(std::log(x.rel / x.rel_prev) / b->inner)
            : 0.0;
    out->sor = sor;
    return 0;
}

}  // extern "C"
