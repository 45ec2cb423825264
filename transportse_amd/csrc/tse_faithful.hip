// tse_faithful.hip -- reference-arithmetic (bit-identical) FP64 kernels.
//
// Compiled with -ffp-contract=off: every source operation is one correctly
// rounded IEEE binary64 operation (division included), in the operand order
// of tse.py, which is what Numba emits with fastmath=False.
//
// Structure of one SOR solve (compute_P / compute_C / compute_T):
//   * assemble: once per call (and again after a dt halving) every cell's
//     coefficients are evaluated exactly as the reference's solve() does and
//     stored as a per-cell record.  Nothing they depend on changes inside the
//     call, so the reference recomputes identical values every sweep; storing
//     them changes no bit of the result and removes ~70 FP64 divisions per
//     cell update.
//   * sweep: a colour half-sweep finishes the reference's update expression
//     from the record and the six neighbours, term by term in its order.
//   * control: one workgroup reduces the change norm and applies the
//     accept / grow / halve / fail logic in a device control block.
#include <algorithm>

#include "tse_internal.h"

namespace tse {

// ---- helpers (tse.py:248-310) -------------------------------------------------

__device__ inline const double *spacing(const Geo &G, int n) {
    return n == 0 ? G.dx : (n == 1 ? G.dy : G.dz);
}

__device__ inline double dsp(const Geo &G, int n, int i, int j, int k) {
    return n == 0 ? G.dx[i] : (n == 1 ? G.dy[j] : G.dz[k]);
}

// harm (tse.py:297-310)
__device__ inline double harm(const Geo &G, const double *H, int i1, int j1,
                              int k1, int i2, int j2, int k2, int n) {
    const double h1 = H[slot(G, i1, j1, k1)], h2 = H[slot(G, i2, j2, k2)];
    if (h1 != 0 && h2 != 0) {
        const double d1 = dsp(G, n, i1, j1, k1), d2 = dsp(G, n, i2, j2, k2);
        return (d1 + d2) / (d1 / h1 + d2 / h2);
    }
    return 0.0;
}

// amean (tse.py:284-294)
__device__ inline double amean(const Geo &G, const double *X, int i1, int j1,
                               int k1, int i2, int j2, int k2, int n) {
    const double x1 = X[slot(G, i1, j1, k1)], x2 = X[slot(G, i2, j2, k2)];
    const double d1 = dsp(G, n, i1, j1, k1), d2 = dsp(G, n, i2, j2, k2);
    return (x1 * d1 + x2 * d2) / (d1 + d2);
}

// alpha (tse.py:249-260); d1/d2 index the spacing array of axis n, d2 is the
// centre cell.
__device__ inline double alpha(const Geo &G, int i1, int j1, int k1, int i2,
                               int j2, int k2, const double *nue,
                               const double *rho, const double *K, int d1,
                               int d2, int n) {
    const double *d = spacing(G, n);
    const double hk = harm(G, K, i1, j1, k1, i2, j2, k2, n);
    const double hn = harm(G, nue, i1, j1, k1, i2, j2, k2, n);
    const double hr = harm(G, rho, i1, j1, k1, i2, j2, k2, n);
    return hk / hn / (d[d2] * (d[d1] + d[d2]) * 0.5) * hr;
}

// beta_g (tse.py:263-271)
__device__ inline double beta_g(const Geo &G, int i1, int j1, int k1, int i2,
                                int j2, int k2, const double *rho, int d) {
    if (d == 0)
        return amean(G, rho, i1, j1, k1, i2, j2, k2, 0) * G.g * 0.5 *
               (G.dx[i1] + G.dx[i2]) * G.sinb;
    return amean(G, rho, i1, j1, k1, i2, j2, k2, 2) * G.g * 0.5 *
           (G.dz[k1] + G.dz[k2]) * G.cosb;
}

// Python's min(q, 0) / max(q, 0) (first argument wins ties).
__device__ inline double py_min0(double q) { return (0.0 < q) ? 0.0 : q; }
__device__ inline double py_max0(double q) { return (0.0 > q) ? 0.0 : q; }

__device__ inline double tavg(double q1, double q2, bool edge_ok, double q3,
                              double q4) {
    if (edge_ok) return (q1 + q2 + q3 + q4) / 4.0;
    return (q1 + q2 + 0.0 + 0.0) / 2.0;
}

// dispersivity (tse.py:460-644): out = D_xl, D_yl, D_zl, D_xr, D_yr, D_zr.
__device__ void dispersivity(const Geo &G, const double *qx, const double *qy,
                             const double *qz, const double *l_l,
                             const double *l_t, const double *B, int i, int j,
                             int k, double out[6]) {
    const int nx = G.nx, ny = G.ny, nz = G.nz;
#define Q(a, ii, jj, kk) (a)[slot(G, ii, jj, kk)]
    double qxl = 0, qyl = 0, qzl = 0, qxr = 0, qyr = 0, qzr = 0;
    double qxyl = 0, qxzl = 0, qyxl = 0, qyzl = 0, qzxl = 0, qzyl = 0;
    double qxyr = 0, qxzr = 0, qyxr = 0, qyzr = 0, qzxr = 0, qzyr = 0;
    if (i != 0) {
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
    if (i != nx - 1) {
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
    if (j != 0) {
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
    if (j != ny - 1) {
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
    if (k != 0) {
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
    if (k != nz - 1) {
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
    out[0] = ll * fabs(qxl) + lt * (fabs(qxyl) + fabs(qxzl));
    out[1] = ll * fabs(qyl) + lt * (fabs(qyxl) + fabs(qyzl));
    out[2] = ll * fabs(qzl) + lt * (fabs(qzxl) + fabs(qzyl));
    out[3] = ll * fabs(qxr) + lt * (fabs(qxyr) + fabs(qxzr));
    out[4] = ll * fabs(qyr) + lt * (fabs(qyxr) + fabs(qyzr));
    out[5] = ll * fabs(qzr) + lt * (fabs(qzxr) + fabs(qzyr));
}

// ---- block reduction of two FP64 sums ----------------------------------------------

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums (a, b) over a block of NT threads; thread 0 writes them to out[0..1].
template <int NT = kBlock>
__device__ inline void block_sum2(double a, double b, double *out) {
    __shared__ double sa[NT / 64], sb[NT / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
#pragma unroll
        for (int q = 0; q < NT / 64; ++q) {
            ta += sa[q];
            tb += sb[q];
        }
        out[0] = ta;
        out[1] = tb;
    }
}

// Zero the record and mask of the pad slots (r = S in each colour when S is
// odd): a pair sweep reads them with the slot before and must see mask 0.
__device__ inline void zero_pad_records(const Geo &G, int nfields, double *rec,
                                        uint8_t *mask) {
    if (G.Sp == G.S || blockIdx.x != 0 || threadIdx.x >= 2) return;
    const long ri = (long)threadIdx.x * G.Sp + G.S;
    for (int f = 0; f < nfields; ++f) rec[f * G.Fr + ri] = 0.0;
    mask[ri] = 0;
}

// ---- compute_P records ---------------------------------------------------------------
// Record fields (each a 2*S array): A_, C_, E_, G_, I_, K_ (face coefficients
// toward x+, x-, y+, y-, z+, z-), A_*B_, C_*D_, I_*J_, K_*L_ (gravity
// products), R*Pn, WP, R+M.

__global__ void __launch_bounds__(kBlock)
p_assemble_kernel(Geo G, const double *__restrict__ Pn,
                  const double *__restrict__ c_f, const double *__restrict__ nue,
                  const double *__restrict__ rho, const double *__restrict__ phi,
                  const double *__restrict__ Kx, const double *__restrict__ Ky,
                  const double *__restrict__ Kz, const double *__restrict__ BP,
                  const double *__restrict__ WP, const double *__restrict__ c_r,
                  double dt, double *__restrict__ rec,
                  uint8_t *__restrict__ mask) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        double v[kPRec];
#pragma unroll
        for (int f = 0; f < kPRec; ++f) v[f] = 0.0;
        uint8_t m = 0;
        if (q.valid) {
            const int i = q.i, j = q.j, k = q.k;
            const int nx = G.nx, ny = G.ny, nz = G.nz;
            double A_ = 0, B_ = 0, C_ = 0, D_ = 0, E_ = 0, G_ = 0, I_ = 0,
                   J_ = 0, K_ = 0, L_ = 0;
            if (i != nx - 1) {
                A_ = alpha(G, i + 1, j, k, i, j, k, nue, rho, Kx, i + 1, i, 0);
                B_ = beta_g(G, i + 1, j, k, i, j, k, rho, 0);
            }
            if (i != 0) {
                C_ = alpha(G, i, j, k, i - 1, j, k, nue, rho, Kx, i - 1, i, 0);
                D_ = beta_g(G, i, j, k, i - 1, j, k, rho, 0);
            }
            if (j != ny - 1)
                E_ = alpha(G, i, j + 1, k, i, j, k, nue, rho, Ky, j + 1, j, 1);
            if (j != 0)
                G_ = alpha(G, i, j, k, i, j - 1, k, nue, rho, Ky, j - 1, j, 1);
            if (k != nz - 1) {
                I_ = alpha(G, i, j, k + 1, i, j, k, nue, rho, Kz, k + 1, k, 2);
                J_ = beta_g(G, i, j, k + 1, i, j, k, rho, 2);
            }
            if (k != 0) {
                K_ = alpha(G, i, j, k, i, j, k - 1, nue, rho, Kz, k - 1, k, 2);
                L_ = beta_g(G, i, j, k, i, j, k - 1, rho, 2);
            }
            const double por = phi[t];
            const double R = rho[t] * ((1 - por) * c_r[t] + por * c_f[t]) / dt;
            const double M = A_ + C_ + E_ + G_ + I_ + K_;
            v[0] = A_;
            v[1] = C_;
            v[2] = E_;
            v[3] = G_;
            v[4] = I_;
            v[5] = K_;
            v[6] = A_ * B_;
            v[7] = C_ * D_;
            v[8] = I_ * J_;
            v[9] = K_ * L_;
            v[10] = R * Pn[t];
            v[11] = WP[t];
            v[12] = R + M;
            m = (BP[t] != 0) ? 1 : 3;  // bit0 valid cell, bit1 updated
        }
        const long ri = t + (long)c * (G.Sp - G.S);  // record slot
#pragma unroll
        for (int f = 0; f < kPRec; ++f) rec[f * G.Fr + ri] = v[f];
        mask[ri] = m;
    }
    zero_pad_records(G, kPRec, rec, mask);
}

// The six other-colour neighbours of a cell (0 outside the grid).  Every
// load is issued, at a slot clamped into the array, and the outside ones are
// replaced by 0 afterwards, so a cell's loads go out together instead of one
// branch at a time.
struct Nb6 {
    double xp, xm, yp, ym, zp, zm;
};

__device__ inline Nb6 neighbours6(const Geo &G, const Cell &q,
                                  const double *__restrict__ Xo) {
    const long rowb = (long)q.row * G.nxh + q.ih;
    const long last = (long)G.S - 1, plane = (long)G.ny * G.nxh;
    auto ld = [&](long o) { return Xo[o < 0 ? 0 : (o > last ? last : o)]; };
    const double xp = ld(rowb + q.s), xm = ld(rowb + q.s - 1);
    const double yp = ld(rowb + G.nxh), ym = ld(rowb - G.nxh);
    const double zp = ld(rowb + plane), zm = ld(rowb - plane);
    Nb6 n;
    n.xp = q.i != G.nx - 1 ? xp : 0.0;
    n.xm = q.i != 0 ? xm : 0.0;
    n.yp = q.j != G.ny - 1 ? yp : 0.0;
    n.ym = q.j != 0 ? ym : 0.0;
    n.zp = q.k != G.nz - 1 ? zp : 0.0;
    n.zm = q.k != 0 ? zm : 0.0;
    return n;
}

// Control step after an iteration (tse.py:718-740 and the transport
// counterparts 896-914, 1064-1079): `C` is the state before the iteration,
// (sd, so) the norm sums of its two half-sweeps (used on check iterations).
// A stopped solve keeps its state.
__device__ inline Ctl control_apply(Ctl C, bool check, double sd, double so) {
    if (C.status != RUNNING) return C;
    C.num_it += 1;
    C.num_sub_it += 1;
    double diff = 1e100;
    if (check) {
        diff = sqrt(sd) / sqrt(so + 1e-200);
        if (isnan(diff)) diff = 1.0;
    }
    C.diff = diff;
    if (diff < C.conv || C.omega_zero) {
        C.t += C.dt;
        if ((double)C.num_sub_it < C.max_it / 2.0) {
            const double g = C.dt * C.inc_it;
            C.dt = (C.max_dt < g) ? C.max_dt : g;
        }
        C.status = CONVERGED;
    } else if (C.num_sub_it >= C.max_it) {
        C.status = NEED_HALVE;  // the host halves dt and restores the field
    } else if (C.dt <= C.min_dt) {
        C.status = FAILED;
    }
    return C;
}

// Norm partials of one iteration: block b of colour c wrote its two sums at
// partials[2*(c*kMaxSweepBlocks + b)].
__device__ inline void partial_pair(const double *__restrict__ partials, int nblocks,
                                    int b, double &sd, double &so) {
    const int colour = b >= nblocks ? 1 : 0;
    const double *p = partials + 2 * ((long)colour * kMaxSweepBlocks + (b - colour * nblocks));
    sd += p[0];
    so += p[1];
}

// Stand-alone control step (one block of kCtlBlock threads): state `in`,
// the iteration's partials; writes the new state to both slots ctl[0..1].
__global__ void __launch_bounds__(kCtlBlock)
control_kernel(const Ctl *__restrict__ in, Ctl *__restrict__ ctl,
               const double *__restrict__ partials, int nblocks) {
    __shared__ int s_check;
    if (threadIdx.x == 0)
        s_check = in->status == RUNNING && is_check_iter(in->check_mode, in->num_it + 1);
    __syncthreads();
    double sd = 0.0, so = 0.0;
    if (s_check) {
#pragma unroll 4
        for (int b = threadIdx.x; b < 2 * nblocks; b += blockDim.x)
            partial_pair(partials, nblocks, b, sd, so);
    }
    __shared__ double res[2];
    block_sum2<kCtlBlock>(sd, so, res);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const Ctl r = control_apply(*in, s_check != 0, res[0], res[1]);
    ctl[0] = r;
    ctl[1] = r;
}

// The control step folded into the next red half-sweep (grids of at most
// kFuseBlocks sweep blocks): every block reads the previous state and the
// previous iteration's partials (one pair per thread, summed in the same
// order in every block, so every block reaches the same state), block 0
// stores the state, and the block sweeps under it.  Returns the state.
__device__ inline Ctl fused_control(const Ctl *__restrict__ in, Ctl *__restrict__ out,
                                    const double *__restrict__ prev, int nblocks) {
    __shared__ Ctl s_ctl;
    __shared__ double res[2];
    double sd = 0.0, so = 0.0;
    if ((int)threadIdx.x < 2 * nblocks) partial_pair(prev, nblocks, threadIdx.x, sd, so);
    block_sum2(sd, so, res);
    if (threadIdx.x == 0) {
        const Ctl c = *in;
        const bool check = is_check_iter(c.check_mode, c.num_it + 1);
        s_ctl = control_apply(c, check, res[0], res[1]);
        if (blockIdx.x == 0) *out = s_ctl;
    }
    __syncthreads();
    return s_ctl;
}

// Half-sweep body shared by compute_P and the transport solvers: each thread
// takes the pair of slots (2m, 2m+1) of the colour, reads both records with
// one 16-byte load per field and the two mask bytes with one load, and
// updates each slot with `update(record, neighbours, old value)`.  All loads
// of a pair are issued before any test; padding slots (mask bit 0 clear) and
// Dirichlet cells (bit 1 clear) discard the result.  With `check`, per-block
// sums of (X_new - X_old)^2 and X_old^2 over valid cells go to partials.
typedef double dpair __attribute__((ext_vector_type(2)));

template <int NREC, typename Update>
__device__ inline void pair_sweep(const Geo &G, int colour, double *__restrict__ X,
                                  const double *__restrict__ rec,
                                  const uint8_t *__restrict__ mask, bool check,
                                  double *__restrict__ partials, Update update) {
    const long rbase = (long)colour * G.Sp, xbase = (long)colour * G.S;
    const double *__restrict__ Xo = X + (long)(1 - colour) * G.S;
    double sd = 0.0, so = 0.0;
    for (int r0 = 2 * (blockIdx.x * blockDim.x + threadIdx.x); r0 < G.S;
         r0 += 2 * gridDim.x * blockDim.x) {
        const long ri = rbase + r0;  // even: the pair is 16-byte aligned
        // the record is read once per half-sweep: non-temporal loads leave
        // the L2 to the other colour's values, which the blocks of the
        // neighbouring rows and planes read again (256x256x128 compute_P
        // half-sweep 99.3 -> 90.5 us)
        dpair R[NREC];
#pragma unroll
        for (int f = 0; f < NREC; ++f)
            R[f] = __builtin_nontemporal_load(
                reinterpret_cast<const dpair *>(rec + f * G.Fr + ri));
        const uint16_t mm = *reinterpret_cast<const uint16_t *>(mask + ri);
        // r0 = S-1 (S odd): the second slot is the pad slot, mask 0
        const int r1 = r0 + 1 < G.S ? r0 + 1 : r0;
        const double X0 = X[xbase + r0], X1 = X[xbase + r1];
        const Nb6 n0 = neighbours6(G, decode(G, colour, r0), Xo);
        const Nb6 n1 = neighbours6(G, decode(G, colour, r1), Xo);
        double a[NREC], b[NREC];
#pragma unroll
        for (int f = 0; f < NREC; ++f) {
            a[f] = R[f].x;
            b[f] = R[f].y;
        }
        const double u0 = update(a, n0, X0), u1 = update(b, n1, X1);
        const unsigned m0 = mm & 0xffu, m1 = mm >> 8;
        const bool up0 = (m0 & 3) == 3, up1 = (m1 & 3) == 3;
        const double N0 = up0 ? u0 : X0, N1 = up1 ? u1 : X1;
        if (up0) X[xbase + r0] = N0;
        if (up1) X[xbase + r1] = N1;
        if (check) {
            if (m0 & 1) {
                const double d = N0 - X0;
                sd += d * d;
                so += X0 * X0;
            }
            if (m1 & 1) {
                const double d = N1 - X1;
                sd += d * d;
                so += X1 * X1;
            }
        }
    }
    if (check)
        block_sum2(sd, so,
                   partials + 2 * ((long)colour * kMaxSweepBlocks + blockIdx.x));
}

// One colour half-sweep of compute_P (tse.py:705-716), update expression of
// tse.py:692-695 evaluated left to right.  If the iteration checks the norm,
// per-block sums of (P_new-P_old)^2 and P_old^2 go to partials.
__global__ void __launch_bounds__(kBlock)
p_sweep_kernel(Geo G, int colour, double *__restrict__ P,
               const double *__restrict__ rec, const uint8_t *__restrict__ mask,
               double sor, SweepCtl sc) {
    bool check = false;
    if (sc.in) {
        const Ctl c = sc.fuse ? fused_control(sc.in, sc.out, sc.prev, sc.nblocks)
                              : *sc.in;
        if (c.status != RUNNING) return;
        check = is_check_iter(c.check_mode, c.num_it + 1);
    }
    double *__restrict__ partials = sc.partials;
    pair_sweep<kPRec>(G, colour, P, rec, mask, check, partials,
                      [sor](const double *rr, const Nb6 &n, double Pc) {
                          const double Anew = rr[0] * n.xp;
                          const double Cnew = rr[1] * n.xm;
                          const double Enew = rr[2] * n.yp;
                          const double Gnew = rr[3] * n.ym;
                          const double Inew = rr[4] * n.zp;
                          const double Knew = rr[5] * n.zm;
                          double s = Anew + rr[6];
                          s = s + Cnew;
                          s = s - rr[7];
                          s = s + Enew;
                          s = s + Gnew;
                          s = s + Inew;
                          s = s + rr[8];
                          s = s + Knew;
                          s = s - rr[9];
                          s = s + rr[10];
                          s = s + rr[11];
                          return (1 - sor) * Pc + sor * s / rr[12];
                      });
}

// ---- compute_q (tse.py:744-780) --------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
q_kernel(Geo G, const double *__restrict__ P, const double *__restrict__ BP,
         const double *__restrict__ Kx, const double *__restrict__ Ky,
         const double *__restrict__ Kz, double *__restrict__ qx,
         double *__restrict__ qy, double *__restrict__ qz,
         const double *__restrict__ nue, const double *__restrict__ rho) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        if (!q.valid) continue;
        const int i = q.i, j = q.j, k = q.k;
        if (i < G.nx - 1) {
            const int e = slot(G, i + 1, j, k);
            if (!(BP[t] != 0 && BP[e] != 0))
                qx[t] = (-harm(G, Kx, i + 1, j, k, i, j, k, 0) /
                         harm(G, nue, i + 1, j, k, i, j, k, 0) *
                         ((P[e] - P[t]) / (0.5 * (G.dx[i + 1] + G.dx[i])) +
                          amean(G, rho, i + 1, j, k, i, j, k, 0) * G.g * G.sinb));
            else
                qx[t] = 0;
        }
        if (j < G.ny - 1) {
            const int n = slot(G, i, j + 1, k);
            if (!(BP[t] != 0 && BP[n] != 0))
                qy[t] = (-harm(G, Ky, i, j + 1, k, i, j, k, 1) /
                         harm(G, nue, i, j + 1, k, i, j, k, 1) * (P[n] - P[t]) /
                         (0.5 * (G.dy[j + 1] + G.dy[j])));
            else
                qy[t] = 0;
        }
        if (k < G.nz - 1) {
            const int u = slot(G, i, j, k + 1);
            if (!(BP[t] != 0 && BP[u] != 0))
                qz[t] = (-harm(G, Kz, i, j, k + 1, i, j, k, 2) /
                         harm(G, nue, i, j, k + 1, i, j, k, 2) *
                         ((P[u] - P[t]) / (0.5 * (G.dz[k + 1] + G.dz[k])) +
                          amean(G, rho, i, j, k + 1, i, j, k, 2) * G.g * G.cosb));
            else
                qz[t] = 0;
        }
    }
}

// ---- smolarkiewicz (tse.py:1083-1116) -------------------------------------------------
// Anti-diffusive velocities: for C > 0, per axis with an upper neighbour,
// Sc*(|q| dd - dt q*q)*(C_nb - C) / ((C + C_nb + eps) dd), left to right; the
// last index of an axis keeps its value; C <= 0 zeroes all three.
__device__ inline double smol_v(double Sc, double q, double dd, double dt,
                                double c0, double c1, double eps) {
    return Sc * (fabs(q) * dd - dt * (q * q)) * (c1 - c0) / ((c0 + c1 + eps) * dd);
}

__global__ void __launch_bounds__(kBlock)
smolarkiewicz_kernel(Geo G, const double *__restrict__ C,
                     const double *__restrict__ qx, const double *__restrict__ qy,
                     const double *__restrict__ qz, double *__restrict__ vx,
                     double *__restrict__ vy, double *__restrict__ vz, double dt,
                     double eps, double Sc) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        if (!q.valid) continue;
        const int i = q.i, j = q.j, k = q.k;
        const double c0 = C[t];
        if (c0 > 0) {
            if (i < G.nx - 1)
                vx[t] = smol_v(Sc, qx[t], 0.5 * (G.dx[i + 1] + G.dx[i]), dt, c0,
                               C[slot(G, i + 1, j, k)], eps);
            if (j < G.ny - 1)
                vy[t] = smol_v(Sc, qy[t], 0.5 * (G.dy[j + 1] + G.dy[j]), dt, c0,
                               C[slot(G, i, j + 1, k)], eps);
            if (k < G.nz - 1)
                vz[t] = smol_v(Sc, qz[t], 0.5 * (G.dz[k + 1] + G.dz[k]), dt, c0,
                               C[slot(G, i, j, k + 1)], eps);
        } else {
            vx[t] = 0.0;
            vy[t] = 0.0;
            vz[t] = 0.0;
        }
    }
}

// ---- flow_error (tse.py:409-446) -------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
flow_error_kernel(Geo G, const double *__restrict__ P,
                  const double *__restrict__ Pn, const double *__restrict__ qx,
                  const double *__restrict__ qy, const double *__restrict__ qz,
                  const double *__restrict__ WP, const double *__restrict__ phi,
                  const double *__restrict__ c_f, const double *__restrict__ c_r,
                  const double *__restrict__ rho, double dt,
                  double *__restrict__ e) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        if (!q.valid) {
            e[t] = 0.0;
            continue;
        }
        const int i = q.i, j = q.j, k = q.k;
        double v = 0.0;
        if (i != G.nx - 1)
            v -= (qx[t] / G.dx[i] * harm(G, rho, i + 1, j, k, i, j, k, 0));
        if (i != 0)
            v += (qx[slot(G, i - 1, j, k)] / G.dx[i] *
                  harm(G, rho, i - 1, j, k, i, j, k, 0));
        if (j != G.ny - 1)
            v -= (qy[t] / G.dy[j] * harm(G, rho, i, j + 1, k, i, j, k, 1));
        if (j != 0)
            v += (qy[slot(G, i, j - 1, k)] / G.dy[j] *
                  harm(G, rho, i, j - 1, k, i, j, k, 1));
        if (k != G.nz - 1)
            v -= (qz[t] / G.dz[k] * harm(G, rho, i, j, k + 1, i, j, k, 2));
        if (k != 0)
            v += (qz[slot(G, i, j, k - 1)] / G.dz[k] *
                  harm(G, rho, i, j, k - 1, i, j, k, 2));
        const double por = phi[t];
        v += (WP[t] + ((1 - por) * c_r[t] + por * c_f[t]) * (P[t] - Pn[t]) *
                          rho[t] / dt);
        e[t] = v;
    }
}

// ---- transport records (compute_C / compute_T) ------------------------------------------
// Record fields: a1..a6 = neighbour coefficients toward x+, x-, y+, y-, z+,
// z- ((A_-N_), (B_+M_), (C_-P_), (D_+O_), (E_-R_), (F_+Q_)), then
// X = (1-omega)*(Xn*M + Mold + ... + Rold), R*Xn, W, E (imbalance term),
// R - omega*M.

// Shared tail of both assemble kernels (tse.py:862-875 / 1026-1043).
__device__ inline void transport_tail(double *v, const double a[6],
                                      const double nb_old[6], double xn,
                                      double R, double M, double W, double EC,
                                      double omega) {
    double so = xn * M;
    so = so + a[0] * nb_old[0];
    so = so + a[1] * nb_old[1];
    so = so + a[2] * nb_old[2];
    so = so + a[3] * nb_old[3];
    so = so + a[4] * nb_old[4];
    so = so + a[5] * nb_old[5];
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = a[f];
    v[6] = (1 - omega) * so;
    v[7] = R * xn;
    v[8] = W;
    v[9] = EC;
    v[10] = R - omega * M;
}

__global__ void __launch_bounds__(kBlock)
c_assemble_kernel(Geo G, const double *__restrict__ Cn,
                  const double *__restrict__ D_fx,
                  const double *__restrict__ D_fy,
                  const double *__restrict__ D_fz,
                  const double *__restrict__ phi, const double *__restrict__ rho,
                  const double *__restrict__ qx, const double *__restrict__ qy,
                  const double *__restrict__ qz, const double *__restrict__ BC,
                  const double *__restrict__ WC, const double *__restrict__ RC,
                  const double *__restrict__ e, const double *__restrict__ lf_l,
                  const double *__restrict__ lf_t, int imblflo, double dt,
                  double omega, double *__restrict__ rec,
                  uint8_t *__restrict__ mask) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        double v[kTRec];
#pragma unroll
        for (int f = 0; f < kTRec; ++f) v[f] = 0.0;
        uint8_t m = 0;
        if (q.valid) {
            const int i = q.i, j = q.j, k = q.k;
            const int nx = G.nx, ny = G.ny, nz = G.nz;
            double D[6] = {0, 0, 0, 0, 0, 0};
            if (lf_l[t] != 0 || lf_t[t] != 0)
                dispersivity(G, qx, qy, qz, lf_l, lf_t, BC, i, j, k, D);
            double A_ = 0, B_ = 0, C_ = 0, D_ = 0, E_ = 0, F_ = 0, G_ = 0,
                   H_ = 0, I_ = 0, J_ = 0, K_ = 0, L_ = 0, M_ = 0, N_ = 0,
                   O_ = 0, P_ = 0, Q_ = 0, R_ = 0;
            double a[6] = {0, 0, 0, 0, 0, 0}, nbo[6] = {0, 0, 0, 0, 0, 0};
            double dd, dd2, ph;
            if (i != nx - 1) {
                const int o = slot(G, i + 1, j, k);
                dd = G.dx[i];
                dd2 = (0.5 * (G.dx[i + 1] + G.dx[i])) * dd;
                ph = amean(G, phi, i + 1, j, k, i, j, k, 0);
                A_ = (harm(G, D_fx, i + 1, j, k, i, j, k, 0) * ph + D[3]) / dd2;
                N_ = py_min0(qx[t]) / dd;
                H_ = py_max0(qx[t]) / dd;
                a[0] = A_ - N_;
                nbo[0] = Cn[o];
            }
            if (i != 0) {
                const int o = slot(G, i - 1, j, k);
                dd = G.dx[i];
                dd2 = (0.5 * (G.dx[i - 1] + G.dx[i])) * dd;
                ph = amean(G, phi, i - 1, j, k, i, j, k, 0);
                B_ = (harm(G, D_fx, i - 1, j, k, i, j, k, 0) * ph + D[0]) / dd2;
                M_ = py_max0(qx[o]) / dd;
                G_ = py_min0(qx[o]) / dd;
                a[1] = B_ + M_;
                nbo[1] = Cn[o];
            }
            if (j != ny - 1) {
                const int o = slot(G, i, j + 1, k);
                dd = G.dy[j];
                dd2 = (0.5 * (G.dy[j + 1] + G.dy[j])) * dd;
                ph = amean(G, phi, i, j + 1, k, i, j, k, 1);
                C_ = (harm(G, D_fy, i, j + 1, k, i, j, k, 1) * ph + D[4]) / dd2;
                P_ = py_min0(qy[t]) / dd;
                J_ = py_max0(qy[t]) / dd;
                a[2] = C_ - P_;
                nbo[2] = Cn[o];
            }
            if (j != 0) {
                const int o = slot(G, i, j - 1, k);
                dd = G.dy[j];
                dd2 = (0.5 * (G.dy[j - 1] + G.dy[j])) * dd;
                ph = amean(G, phi, i, j - 1, k, i, j, k, 1);
                D_ = (harm(G, D_fy, i, j - 1, k, i, j, k, 1) * ph + D[1]) / dd2;
                O_ = py_max0(qy[o]) / dd;
                I_ = py_min0(qy[o]) / dd;
                a[3] = D_ + O_;
                nbo[3] = Cn[o];
            }
            if (k != nz - 1) {
                const int o = slot(G, i, j, k + 1);
                dd = G.dz[k];
                dd2 = (0.5 * (G.dz[k + 1] + G.dz[k])) * dd;
                ph = amean(G, phi, i, j, k + 1, i, j, k, 2);
                E_ = (harm(G, D_fz, i, j, k + 1, i, j, k, 2) * ph + D[5]) / dd2;
                R_ = py_min0(qz[t]) / dd;
                L_ = py_max0(qz[t]) / dd;
                a[4] = E_ - R_;
                nbo[4] = Cn[o];
            }
            if (k != 0) {
                const int o = slot(G, i, j, k - 1);
                dd = G.dz[k];
                dd2 = (0.5 * (G.dz[k - 1] + G.dz[k])) * dd;
                ph = amean(G, phi, i, j, k - 1, i, j, k, 2);
                F_ = (harm(G, D_fz, i, j, k - 1, i, j, k, 2) * ph + D[2]) / dd2;
                Q_ = py_max0(qz[o]) / dd;
                K_ = py_min0(qz[o]) / dd;
                a[5] = F_ + Q_;
                nbo[5] = Cn[o];
            }
            const double R = RC[t] * phi[t] / dt;
            const double M =
                -A_ - B_ - C_ - D_ - E_ - F_ + G_ - H_ + I_ - J_ + K_ - L_;
            const double EC = imblflo ? Cn[t] / rho[t] * e[t] : 0.0;
            transport_tail(v, a, nbo, Cn[t], R, M, WC[t], EC, omega);
            m = (BC[t] != 0) ? 1 : 3;
        }
        const long ri = t + (long)c * (G.Sp - G.S);  // record slot
#pragma unroll
        for (int f = 0; f < kTRec; ++f) rec[f * G.Fr + ri] = v[f];
        mask[ri] = m;
    }
    zero_pad_records(G, kTRec, rec, mask);
}

// One face of compute_T's solve() (tse.py:943-1031).
__device__ inline void t_face(const Geo &G, const double *phi,
                              const double *l_f, const double *l_r,
                              const double *cp_f, const double *rho, int i1,
                              int j1, int k1, int i2, int j2, int k2, int n,
                              double dd, double dd2, double Ddisp, double *A,
                              double *S) {
    const double por = amean(G, phi, i1, j1, k1, i2, j2, k2, n);
    const double lf = harm(G, l_f, i1, j1, k1, i2, j2, k2, n);
    const double lr = harm(G, l_r, i1, j1, k1, i2, j2, k2, n);
    const double cpf = harm(G, cp_f, i1, j1, k1, i2, j2, k2, n);
    const double rhocpf = harm(G, rho, i1, j1, k1, i2, j2, k2, n) * cpf;
    *A = ((1 - por) * lr + por * lf + Ddisp * rhocpf) / dd2;
    *S = rhocpf / dd;
}

__global__ void __launch_bounds__(kBlock)
t_assemble_kernel(Geo G, const double *__restrict__ Tn,
                  const double *__restrict__ l_f,
                  const double *__restrict__ cp_f,
                  const double *__restrict__ rho, const double *__restrict__ phi,
                  const double *__restrict__ qx, const double *__restrict__ qy,
                  const double *__restrict__ qz, const double *__restrict__ BT,
                  const double *__restrict__ WT, const double *__restrict__ e,
                  const double *__restrict__ l_r,
                  const double *__restrict__ cp_r,
                  const double *__restrict__ rho_r,
                  const double *__restrict__ lt_l,
                  const double *__restrict__ lt_t, int imblflo, double dt,
                  double omega, double *__restrict__ rec,
                  uint8_t *__restrict__ mask) {
    const long F = 2L * G.S;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const int c = t >= G.S ? 1 : 0;
        const Cell q = decode(G, c, (int)(t - (long)c * G.S));
        double v[kTRec];
#pragma unroll
        for (int f = 0; f < kTRec; ++f) v[f] = 0.0;
        uint8_t m = 0;
        if (q.valid) {
            const int i = q.i, j = q.j, k = q.k;
            const int nx = G.nx, ny = G.ny, nz = G.nz;
            double D[6] = {0, 0, 0, 0, 0, 0};
            if (lt_l[t] != 0 || lt_t[t] != 0)
                dispersivity(G, qx, qy, qz, lt_l, lt_t, BT, i, j, k, D);
            double A_ = 0, B_ = 0, C_ = 0, D_ = 0, E_ = 0, F_ = 0, G_ = 0,
                   H_ = 0, I_ = 0, J_ = 0, K_ = 0, L_ = 0, S_ = 0, M_ = 0,
                   N_ = 0, O_ = 0, P_ = 0, Q_ = 0, R_ = 0;
            double a[6] = {0, 0, 0, 0, 0, 0}, nbo[6] = {0, 0, 0, 0, 0, 0};
            if (i != nx - 1) {
                const int o = slot(G, i + 1, j, k);
                const double dd = G.dx[i];
                const double dd2 = (0.5 * (G.dx[i + 1] + G.dx[i])) * dd;
                t_face(G, phi, l_f, l_r, cp_f, rho, i, j, k, i + 1, j, k, 0, dd,
                       dd2, D[3], &A_, &S_);
                N_ = py_min0(qx[t]) * S_;
                H_ = py_max0(qx[t]) * S_;
                a[0] = A_ - N_;
                nbo[0] = Tn[o];
            }
            if (i != 0) {
                const int o = slot(G, i - 1, j, k);
                const double dd = G.dx[i];
                const double dd2 = (0.5 * (G.dx[i - 1] + G.dx[i])) * dd;
                t_face(G, phi, l_f, l_r, cp_f, rho, i, j, k, i - 1, j, k, 0, dd,
                       dd2, D[0], &B_, &S_);
                M_ = py_max0(qx[o]) * S_;
                G_ = py_min0(qx[o]) * S_;
                a[1] = B_ + M_;
                nbo[1] = Tn[o];
            }
            if (j != ny - 1) {
                const int o = slot(G, i, j + 1, k);
                const double dd = G.dy[j];
                const double dd2 = (0.5 * (G.dy[j + 1] + G.dy[j])) * dd;
                t_face(G, phi, l_f, l_r, cp_f, rho, i, j + 1, k, i, j, k, 1, dd,
                       dd2, D[4], &C_, &S_);
                P_ = py_min0(qy[t]) * S_;
                J_ = py_max0(qy[t]) * S_;
                a[2] = C_ - P_;
                nbo[2] = Tn[o];
            }
            if (j != 0) {
                const int o = slot(G, i, j - 1, k);
                const double dd = G.dy[j];
                const double dd2 = (0.5 * (G.dy[j - 1] + G.dy[j])) * dd;
                t_face(G, phi, l_f, l_r, cp_f, rho, i, j - 1, k, i, j, k, 1, dd,
                       dd2, D[1], &D_, &S_);
                O_ = py_max0(qy[o]) * S_;
                I_ = py_min0(qy[o]) * S_;
                a[3] = D_ + O_;
                nbo[3] = Tn[o];
            }
            if (k != nz - 1) {
                const int o = slot(G, i, j, k + 1);
                const double dd = G.dz[k];
                const double dd2 = (0.5 * (G.dz[k + 1] + G.dz[k])) * dd;
                t_face(G, phi, l_f, l_r, cp_f, rho, i, j, k + 1, i, j, k, 2, dd,
                       dd2, D[5], &E_, &S_);
                R_ = py_min0(qz[t]) * S_;
                L_ = py_max0(qz[t]) * S_;
                a[4] = E_ - R_;
                nbo[4] = Tn[o];
            }
            if (k != 0) {
                const int o = slot(G, i, j, k - 1);
                const double dd = G.dz[k];
                const double dd2 = (0.5 * (G.dz[k - 1] + G.dz[k])) * dd;
                t_face(G, phi, l_f, l_r, cp_f, rho, i, j, k - 1, i, j, k, 2, dd,
                       dd2, D[2], &F_, &S_);
                Q_ = py_max0(qz[o]) * S_;
                K_ = py_min0(qz[o]) * S_;
                a[5] = F_ + Q_;
                nbo[5] = Tn[o];
            }
            const double por_ = phi[t];
            const double R = (por_ * rho[t] * cp_f[t] +
                              (1 - por_) * rho_r[t] * cp_r[t]) /
                             dt;
            const double M =
                -A_ - B_ - C_ - D_ - E_ - F_ + G_ - H_ + I_ - J_ + K_ - L_;
            const double ET = imblflo ? cp_f[t] * Tn[t] * e[t] : 0.0;
            transport_tail(v, a, nbo, Tn[t], R, M, WT[t], ET, omega);
            m = (BT[t] != 0) ? 1 : 3;
        }
        const long ri = t + (long)c * (G.Sp - G.S);  // record slot
#pragma unroll
        for (int f = 0; f < kTRec; ++f) rec[f * G.Fr + ri] = v[f];
        mask[ri] = m;
    }
    zero_pad_records(G, kTRec, rec, mask);
}

// One colour half-sweep of compute_C / compute_T (tse.py:883-894,
// 1051-1062), update expression of tse.py:870-875 / 1038-1043.
__global__ void __launch_bounds__(kBlock)
tr_sweep_kernel(Geo G, int colour, double *__restrict__ X,
                const double *__restrict__ rec,
                const uint8_t *__restrict__ mask, double sor, double omega,
                SweepCtl sc) {
    bool check = false;
    if (sc.in) {
        const Ctl c = sc.fuse ? fused_control(sc.in, sc.out, sc.prev, sc.nblocks)
                              : *sc.in;
        if (c.status != RUNNING) return;
        check = is_check_iter(c.check_mode, c.num_it + 1);
    }
    double *__restrict__ partials = sc.partials;
    pair_sweep<kTRec>(G, colour, X, rec, mask, check, partials,
                      [sor, omega](const double *rr, const Nb6 &n, double Xc) {
                          double s = rr[0] * n.xp;
                          s = s + rr[1] * n.xm;
                          s = s + rr[2] * n.yp;
                          s = s + rr[3] * n.ym;
                          s = s + rr[4] * n.zp;
                          s = s + rr[5] * n.zm;
                          double u = omega * s + rr[6];
                          u = u + rr[7];
                          u = u + rr[8];
                          u = u - rr[9];
                          return (1 - sor) * Xc + sor * u / rr[10];
                      });
}

// ---- l2_norm of two fields --------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
l2_partial_kernel(Geo G, const double *__restrict__ New,
                  const double *__restrict__ Old, double *__restrict__ partials) {
    const long F = 2L * G.S;
    double sd = 0.0, so = 0.0;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < F;
         t += (long)gridDim.x * blockDim.x) {
        const double d = New[t] - Old[t];
        sd += d * d;
        so += Old[t] * Old[t];
    }
    block_sum2(sd, so, partials + 2 * blockIdx.x);
}

__global__ void __launch_bounds__(kBlock)
l2_final_kernel(const double *__restrict__ partials, int nblocks,
                double *__restrict__ out) {
    double sd = 0.0, so = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        sd += partials[2 * b];
        so += partials[2 * b + 1];
    }
    __shared__ double res[2];
    block_sum2(sd, so, res);
    __syncthreads();
    if (threadIdx.x == 0) {
        double l2 = sqrt(res[0]) / sqrt(res[1] + 1e-200);
        out[0] = isnan(l2) ? 1.0 : l2;
        out[1] = isnan(l2) ? 1.0 : 0.0;
    }
}

// ---- launchers ------------------------------------------------------------------------------

static inline int grid_for(long n) {
    long b = (n + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

int launch_p_assemble(tse_ctx *ctx, const tse_p_fields *f, double dt) {
    const Geo &G = ctx->geo;
    const long F = 2L * G.S;
    int rc = ensure_rec(ctx, rec_buffer_elems(G, kPRec));
    if (rc) return rc;
    rc = ensure_mask(ctx, mask_bytes(G));
    if (rc) return rc;
    hipLaunchKernelGGL(p_assemble_kernel, dim3(grid_for(F)), dim3(kBlock), 0,
                       ctx->stream, G, f->Pn, f->c_f, f->nue, f->rho_f, f->phi,
                       f->Kx, f->Ky, f->Kz, f->BP, f->WP, f->c_r, dt, ctx->rec,
                       ctx->mask);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_p_sweep(tse_ctx *ctx, double *P, int colour, double sor,
                   const SweepCtl &sc) {
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(p_sweep_kernel, dim3(pair_sweep_blocks(G)), dim3(kBlock), 0,
                       ctx->stream, G, colour, P, ctx->rec, ctx->mask, sor, sc);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_control(tse_ctx *ctx, const Ctl *in, const double *partials,
                   int nblocks) {
    hipLaunchKernelGGL(control_kernel, dim3(1), dim3(kCtlBlock), 0, ctx->stream,
                       in, ctx->ctl, partials, nblocks);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_q(tse_ctx *ctx, const double *P, const double *BP,
             const double *Kx, const double *Ky, const double *Kz,
             double *qx, double *qy, double *qz, const double *nue,
             const double *rho_f) {
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(q_kernel, dim3(grid_for(2L * G.S)), dim3(kBlock), 0,
                       ctx->stream, G, P, BP, Kx, Ky, Kz, qx, qy, qz, nue,
                       rho_f);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_smolarkiewicz(tse_ctx *ctx, const double *C, const double *qx,
                         const double *qy, const double *qz, double *vx,
                         double *vy, double *vz, double dt, double eps,
                         double Sc) {
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(smolarkiewicz_kernel, dim3(grid_for(2L * G.S)),
                       dim3(kBlock), 0, ctx->stream, G, C, qx, qy, qz, vx, vy,
                       vz, dt, eps, Sc);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_flow_error(tse_ctx *ctx, const double *P, const double *Pn,
                      const double *qx, const double *qy, const double *qz,
                      const double *WP, const double *phi, const double *c_f,
                      const double *c_r, const double *rho_f, double dt,
                      double *e) {
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(flow_error_kernel, dim3(grid_for(2L * G.S)),
                       dim3(kBlock), 0, ctx->stream, G, P, Pn, qx, qy, qz, WP,
                       phi, c_f, c_r, rho_f, dt, e);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_c_assemble(tse_ctx *ctx, const tse_c_fields *f, double dt,
                      double omega) {
    const Geo &G = ctx->geo;
    const long F = 2L * G.S;
    int rc = ensure_rec(ctx, rec_buffer_elems(G, kTRec));
    if (rc) return rc;
    rc = ensure_mask(ctx, mask_bytes(G));
    if (rc) return rc;
    hipLaunchKernelGGL(c_assemble_kernel, dim3(grid_for(F)), dim3(kBlock), 0,
                       ctx->stream, G, f->Cn, f->D_fx, f->D_fy, f->D_fz, f->phi,
                       f->rho_f, f->qx, f->qy, f->qz, f->BC, f->WC, f->RC, f->e,
                       f->lf_l, f->lf_t, (int)f->imblflo, dt, omega, ctx->rec,
                       ctx->mask);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_t_assemble(tse_ctx *ctx, const tse_t_fields *f, double dt,
                      double omega) {
    const Geo &G = ctx->geo;
    const long F = 2L * G.S;
    int rc = ensure_rec(ctx, rec_buffer_elems(G, kTRec));
    if (rc) return rc;
    rc = ensure_mask(ctx, mask_bytes(G));
    if (rc) return rc;
    hipLaunchKernelGGL(t_assemble_kernel, dim3(grid_for(F)), dim3(kBlock), 0,
                       ctx->stream, G, f->Tn, f->l_f, f->cp_f, f->rho_f, f->phi,
                       f->qx, f->qy, f->qz, f->BT, f->WT, f->e, f->l_r, f->cp_r,
                       f->rho_r, f->lt_l, f->lt_t, (int)f->imblflo, dt, omega,
                       ctx->rec, ctx->mask);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_tr_sweep(tse_ctx *ctx, double *X, int colour, double sor,
                    double omega, const SweepCtl &sc) {
    const Geo &G = ctx->geo;
    hipLaunchKernelGGL(tr_sweep_kernel, dim3(pair_sweep_blocks(G)), dim3(kBlock),
                       0, ctx->stream, G, colour, X, ctx->rec, ctx->mask, sor,
                       omega, sc);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

int launch_l2(tse_ctx *ctx, const double *New, const double *Old) {
    const Geo &G = ctx->geo;
    const int nb = std::min(sweep_blocks(2 * G.S), 4096);
    hipLaunchKernelGGL(l2_partial_kernel, dim3(nb), dim3(kBlock), 0,
                       ctx->stream, G, New, Old, ctx->partials);
    TSE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(l2_final_kernel, dim3(1), dim3(kBlock), 0, ctx->stream,
                       ctx->partials, nb, ctx->red_out);
    TSE_HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace tse
