// Precision form of the fluid step, and its FP64 twin.
//
// No counterpart in the reference, whose Jacobi loops run a fixed count in
// FP32 (src/fluid.cpp:151,239).  The precision form keeps the reference's
// step structure (src/fluid.cpp:298-310) and changes three things:
//
//   - the 5-term stencil sums of both Jacobi sweeps (src/fluid.cpp:19-21)
//     are compensated: Kahan's compensated summation of the neighbours,
//     the product by alpha with its exact FMA error, the diffusion's centre
//     term by TwoSum, so each numerator is rounded once from about twice FP32
//     precision before the division;
//   - every pressure sweep also accumulates, in FP64, the residual of its
//     input iterate, r = b + pL + pR + pT + pB - 4p (zero at the fixed point
//     of the sweep p' = (pL+pR+pT+pB+b)/4), from the compensated sum, and
//     sum b^2 once;
//   - the pressure iteration stops at the first iterate i >= 1 with
//     ||r(p^i)|| <= tol ||b||, at most K sweeps, decided on the device.
//
// Kernels (temporally blocked like the plain step, pfs_step.hip, on the
// same register-strip tiles, pfs_step_kernels.h):
//
//   k_diffuse_precise     advect + C diffusion sweeps per launch (compensated),
//                         the last launch also the divergence and the
//                         pressure start (diffuse_tile<..., COMP = true>);
//                         the first also clears the step's control block
//   k_pressure_precise    C pressure sweeps per launch from an iterate plane,
//                         with the FP64 residual of every iterate the launch
//                         reads on the tile's own cells; launch `launch`
//                         covers iterates launch*C .. launch*C + C.
//
// Every workgroup writes its per-row FP64 sums to its own slot (one per
// tile); the last workgroup of a launch to finish folds each row over the
// tiles in tile order (a fixed order: the totals, and so the early exit, do
// not depend on scheduling) and decides.  The next launch reads that
// decision at entry:
//   - no iterate converged, more sweeps to go: continue from the plane the
//     last launch wrote (planes rotate by launch: launch k reads k % 3 and
//     writes (k+1) % 3);
//   - iterate i converged: re-run the sweeps that end at i from the plane
//     that precedes them, still intact, then the gradient subtraction and the
//     outputs; later launches return at entry.
// The last chunk writes the outputs for Kp = K as it goes; one more launch
// fixes them up when the fold of that chunk finds an earlier converged
// iterate.  So the early exit costs no host round trip.
//
// Buffer contract, with Kd = K the diffusion sweeps and Kp <= K the pressure
// sweeps of the result: E = D^Kd if Kd is even else D^(Kd-1); the pressure
// starts from tmp[2] if Kd is even else vp[2];
//   tmp = (E.u, E.v, p^Kp, b)     vp = (E - grad p^Kp, p^(Kp-1), b)
// which is exactly the reference's step when Kp = Kd (the plain step's
// contract, see pfs_step.hip).
//
// The FP64 twin runs the same contract in double with plain arithmetic, one
// launch per sweep and given (Kd, Kp): the yardstick the precision form's
// error is reported against.  With Kd = Kp = K it reproduces the
// reference's algorithm in double exactly.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#include "pfs.h"
#include "pfs_internal.h"
#include "pfs_step_kernels.h"

namespace pfs {
namespace {

constexpr int kMaxK = PFS_PRECISE_MAX_ITERS;

// ============================ FP64 twin ======================================
constexpr int kBX = 64, kBY = 4;
constexpr int kRows = 8;  // rows per thread: a block covers kBX x (kBY * kRows) cells

struct Cell {
    int i, j, im, ip, jm, jp;
    __device__ Cell(int w, int h, int i_, int j_) : i(i_), j(j_) {
        im = wrap_minus(i, w);
        ip = wrap_plus(i, w);
        jm = wrap_minus(j, h);
        jp = wrap_plus(j, h);
    }
};

// Calls f(cell) for each of this thread's cells: column blockIdx.x*kBX + x,
// rows kRows apart by kBY within the block's kBY*kRows rows.
template <typename F>
__device__ __forceinline__ void for_cells(int w, int h, F &&f) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j0 = blockIdx.y * (kBY * kRows) + threadIdx.y;
    if (i >= w) return;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        const int j = j0 + r * kBY;
        if (j < h) f(Cell(w, h, i, j));
    }
}

// Advection of vp[0,1] (src/fluid.cpp:24-70) into the iterate plane d.
__global__ __launch_bounds__(kBX *kBY) void k_t_advect(const double *__restrict__ P,
                                                      double2 *__restrict__ d, int w, int h,
                                                      double dt, double fw, double fh,
                                                      double inv_w, double inv_h, int pow2_w,
                                                      int pow2_h) {
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const double xp = (double)c.i - div_n(dt * P[4 * cell], fw, inv_w, pow2_w);
        const double yp = (double)c.j - div_n(dt * P[4 * cell + 1], fh, inv_h, pow2_h);
        const TraceT<double> t = trace(xp, yp, w, h, fw, fh);
        const int a00 = 4 * (t.j0 * w + t.i0), a10 = 4 * (t.j0 * w + t.i1);
        const int a01 = 4 * (t.j1 * w + t.i0), a11 = 4 * (t.j1 * w + t.i1);
        d[cell] = {bilerp(t.sx, t.sy, P[a00], P[a10], P[a01], P[a11]),
                   bilerp(t.sx, t.sy, P[a00 + 1], P[a10 + 1], P[a01 + 1], P[a11 + 1])};
    });
}

// One diffusion sweep (src/fluid.cpp:154-185) on the iterate planes.
__global__ __launch_bounds__(kBX *kBY) void k_t_diffuse(const double2 *__restrict__ in,
                                                       double2 *__restrict__ out, int w, int h,
                                                       double alpha, double beta) {
    for_cells(w, h, [&](const Cell &c) {
        const double2 L = in[c.j * w + c.im], R = in[c.j * w + c.ip], T = in[c.jm * w + c.i],
                      B = in[c.jp * w + c.i], C = in[c.j * w + c.i];
        out[c.j * w + c.i] = {diffuse_cell(L.x, R.x, T.x, B.x, C.x, alpha, beta),
                              diffuse_cell(L.y, R.y, T.y, B.y, C.y, alpha, beta)};
    });
}

// Divergence of D^Kd (src/fluid.cpp:221-237) into b, and the pressure start
// (tmp[2] for even Kd, vp[2] for odd) into iterate plane 0.
__global__ __launch_bounds__(kBX *kBY) void k_t_divergence(const double2 *__restrict__ dk,
                                                          const double *__restrict__ P,
                                                          const double *__restrict__ Q,
                                                          double *__restrict__ b,
                                                          double *__restrict__ p0, int w, int h,
                                                          double gamma, int kd_even) {
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const double uR = dk[c.j * w + c.ip].x - dk[c.j * w + c.im].x;
        const double vT = dk[c.jp * w + c.i].y - dk[c.jm * w + c.i].y;
        b[cell] = gamma * (uR + vT);
        p0[cell] = kd_even ? Q[4 * cell + 2] : P[4 * cell + 2];
    });
}

// Pressure sweep (src/fluid.cpp:240-257): p^s from p^(s-1).
__global__ __launch_bounds__(kBX *kBY) void k_t_pressure(const double *__restrict__ in,
                                                        double *__restrict__ out,
                                                        const double *__restrict__ b, int w,
                                                        int h) {
    for_cells(w, h, [&](const Cell &c) {
        out[c.j * w + c.i] = pressure_cell(in[c.j * w + c.im], in[c.j * w + c.ip],
                                           in[c.jm * w + c.i], in[c.jp * w + c.i],
                                           b[c.j * w + c.i]);
    });
}

// Gradient subtraction with p^Kp (src/fluid.cpp:269-296) and the output
// contract; iterate k lives in plane k % 3.
__global__ __launch_bounds__(kBX *kBY) void k_t_finalize(
    double *__restrict__ P, double *__restrict__ Q, const double2 *__restrict__ dk,
    const double2 *__restrict__ dkm1, const double *p0, const double *p1, const double *p2,
    const double *__restrict__ b, int w, int h, double dt, int kd, int kp) {
    const double *pk = kp % 3 == 0 ? p0 : kp % 3 == 1 ? p1 : p2;
    const double *pkm1 = kp % 3 == 1 ? p0 : kp % 3 == 2 ? p1 : p2;
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const double pL = pk[c.j * w + c.im], pR = pk[c.j * w + c.ip];
        const double pT = pk[c.jm * w + c.i], pB = pk[c.jp * w + c.i];
        const double gradX = (pR - pL) * dt / 2.0;
        const double gradY = (pB - pT) * dt / 2.0;
        const double2 e = (kd % 2 == 0) ? dk[cell] : dkm1[cell];
        const double bb = b[cell];
        P[4 * cell] = e.x - gradX;
        P[4 * cell + 1] = e.y - gradY;
        P[4 * cell + 2] = pkm1[cell];
        P[4 * cell + 3] = bb;
        Q[4 * cell] = e.x;
        Q[4 * cell + 1] = e.y;
        Q[4 * cell + 2] = pk[cell];
        Q[4 * cell + 3] = bb;
    });
}

// Workspace of the twin: two iterate planes of double2, b and three pressure
// planes of double.
size_t twin_ws_bytes(int w, int h) { return (size_t)w * h * (2 * sizeof(double2) + 4 * sizeof(double)); }

int twin_step(double *P, double *Q, int w, int h, double dt, double visc, int kd, int kp, void *ws,
              hipStream_t s) {
    const size_t N = (size_t)w * h;
    double2 *d[2] = {static_cast<double2 *>(ws), static_cast<double2 *>(ws) + N};
    double *b = reinterpret_cast<double *>(d[1] + N);
    double *p[3] = {b + N, b + 2 * N, b + 3 * N};
    const dim3 grid((w + kBX - 1) / kBX, (h + kBY * kRows - 1) / (kBY * kRows)), block(kBX, kBY);
    const double alpha = visc * dt;
    const double beta = 1.0 + 4.0 * alpha;  // src/fluid.cpp:145
    const double gamma = -1.0 / dt;         // src/fluid.cpp:218
    hipLaunchKernelGGL(k_t_advect, grid, block, 0, s, P, d[0], w, h, dt, (double)w, (double)h,
                       1.0 / w, 1.0 / h, (w & (w - 1)) == 0, (h & (h - 1)) == 0);
    for (int k = 1; k <= kd; ++k)
        hipLaunchKernelGGL(k_t_diffuse, grid, block, 0, s, d[(k - 1) & 1], d[k & 1], w, h, alpha,
                           beta);
    hipLaunchKernelGGL(k_t_divergence, grid, block, 0, s, d[kd & 1], P, Q, b, p[0], w, h, gamma,
                       kd % 2 == 0);
    for (int k = 1; k <= kp; ++k)
        hipLaunchKernelGGL(k_t_pressure, grid, block, 0, s, p[(k - 1) % 3], p[k % 3], b, w, h);
    hipLaunchKernelGGL(k_t_finalize, grid, block, 0, s, P, Q, d[kd & 1], d[(kd + 1) & 1], p[0],
                       p[1], p[2], b, w, h, dt, kd, kp);
    return check_launch("FP64 twin step");
}

// ==================== FP32 precision form (temporally blocked) ===============
constexpr int kRowB2 = kMaxK + 1;      // rows 0..kMaxK: sum r(p^i)^2; this row: sum b^2
constexpr int kNRows = kMaxK + 2;
constexpr int kMaxLaunch = kMaxK + 2;  // pressure launches: chunks of >= 1 sweep, the fix-up

// Device-side state of one precision step (cleared by its first diffusion launch).
struct PCtl {
    unsigned count[kMaxLaunch];  // workgroups finished, per pressure launch
    int done;                    // 1: vp and tmp hold the result
    int stop;                    // first converged iterate a fold found, else -1
    int kp;                      // pressure sweeps of the result
    int pad;
    double total[kNRows];        // folded sums
};

struct PreciseArgs {
    float *P, *Q;               // vp, tmp: the outputs
    float *pl[3];               // pressure iterate planes; pl[0] holds the start
    const float *bpl;           // the divergence b
    const float2 *E;            // the diffused velocity the outputs keep
    double *part;               // per-tile FP64 sums, part[row * ntiles + tile]
    PCtl *ctl;
    int w, h, tiles_x, ntiles;
    float dt;
    int K, C, nchunks, launch;  // sweeps, sweeps per launch, chunks, this launch
    double tol2;
};

// Sum over the 64 lanes of a wave by a fixed butterfly: every lane ends
// with the same value, independent of scheduling.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ctl non-null (the step's first launch): workgroup 0 also clears the
// control block the pressure launches use.
// x plus the value DPP control CTRL moves into this lane (0 where it moves
// none), on the two 32-bit halves of a double.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const int mlo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWS, 0xf, false);
    const int mhi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWS, 0xf, false);
    return x + __hiloint2double(mhi, mlo);
}

// Inclusive scan of x over the wave in a fixed pattern (row_shr 1, 2, 4, 8,
// then row_bcast 15 and 31): lane 63 ends with the wave's sum.  DPP moves
// only, no LDS round trip.
__device__ __forceinline__ double wave_scan_dpp(double x) {
    x = dpp_add<0x111, 0xf>(x);  // row_shr:1
    x = dpp_add<0x112, 0xf>(x);  // row_shr:2
    x = dpp_add<0x114, 0xf>(x);  // row_shr:4
    x = dpp_add<0x118, 0xf>(x);  // row_shr:8
    x = dpp_add<0x142, 0xa>(x);  // row_bcast:15
    x = dpp_add<0x143, 0xc>(x);  // row_bcast:31
    return x;
}

template <int H, bool LAST, int TS>
__global__ __launch_bounds__((Geo<H, TS>::NT)) void k_diffuse_precise(DiffuseArgs a, PCtl *ctl) {
    if (ctl && blockIdx.x == 0) {
        for (int k = threadIdx.x; k < kMaxLaunch; k += blockDim.x) ctl->count[k] = 0u;
        for (int k = threadIdx.x; k < kNRows; k += blockDim.x) ctl->total[k] = 0.0;
        if (threadIdx.x == 0) {
            ctl->done = 0;
            ctl->stop = -1;
            ctl->kp = 0;
        }
    }
    diffuse_tile<H, LAST, TS, true>(a, xcd_tile(blockIdx.x, a.ntiles));
}

// Up to C = H - 1 compensated pressure sweeps of one tile, the residual of
// every iterate they read, and (on the launch that ends the solve) the
// gradient subtraction and the outputs.  See the file comment for the
// launch sequence.
// NTMAX = 512: at most 128 VGPRs, so that two workgroups share a CU.
template <int H, int TS, int NTMAX>
__global__ __launch_bounds__((Geo<H, TS, NTMAX>::NT))
__attribute__((amdgpu_waves_per_eu(NTMAX == 512 ? 4 : 1))) void k_pressure_precise(PreciseArgs a) {
    using G = Geo<H, TS, NTMAX>;
    constexpr int S = G::S, RX = G::RX, RY = G::RY, PL = G::RYP * RX;
    constexpr int C = H - 1;      // most sweeps of one launch
    constexpr int NW = G::NT / 64;
    constexpr int NSLOT = C + 2;  // rows one launch adds: C residuals, r(p^K), b^2
    __shared__ float plane[2 * PL];
    __shared__ double wsum[NSLOT][NW];
    __shared__ double rowsum[NSLOT];
    __shared__ int last_group;

    // ---- what this launch does: every workgroup reads the same state -------
    const int done = a.ctl->done, stop = a.ctl->stop;
    if (done) return;
    bool finish, fin;  // finish: re-run to the converged iterate; fin: write the outputs
    int src, b0, ns;   // input plane, its iterate number, sweeps
    if (stop >= 0) {
        // The fold of launch k = launch - 1 (iterates kC .. kC + C - 1) found
        // iterate `stop`: re-run from launch k's input, or from launch k-1's
        // when `stop` is launch k's input itself (stop >= 1, so k >= 1 then).
        const int k = a.launch - 1, kb = k * a.C;
        finish = fin = true;
        if (stop > kb) {
            src = k % 3;
            b0 = kb;
        } else {
            src = (k + 2) % 3;
            b0 = kb - a.C;
        }
        ns = stop - b0;
    } else if (a.launch < a.nchunks) {
        finish = false;
        src = a.launch % 3;
        b0 = a.launch * a.C;
        ns = min(a.C, a.K - b0);
        fin = a.launch == a.nchunks - 1;
    } else {
        return;  // the fix-up launch, with nothing to fix
    }

    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const Strip<G> st(threadIdx.x, tile, a.tiles_x, a.w, a.h);
    const float *__restrict__ in = a.pl[src];
    float p[S], rhs[S];
    double b2 = 0.0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int g = st.row_ok(k) ? st.gj[k] * a.w + st.gi : 0;
        p[k] = in[g];
        rhs[k] = a.bpl[g];
    }
    if (a.launch == 0) {
#pragma unroll
        for (int k = 0; k < S; ++k) b2 += st.mine(k, a.w, a.h) ? (double)rhs[k] * rhs[k] : 0.0;
    }
    // E, which only the outputs need: loaded before the sweeps so that its
    // latency hides behind them, except in 512-thread workgroups, whose
    // longer strips leave no registers for it.
    constexpr bool kEarlyE = NTMAX == 1024;
    float2 e[S];
    if (kEarlyE && fin) {
#pragma unroll
        for (int k = 0; k < S; ++k) e[k] = a.E[st.mine(k, a.w, a.h) ? st.cell(k, a.w) : 0];
    }

    // Sweeps (src/fluid.cpp:240-257), p' = RN(s - k) / 4 with (s, k) the
    // compensated L + R + T + B + b; the residual of the iterate read,
    // r = (s - 4p) - k in FP64, on the tile's own cells, summed over the
    // thread's rows, then over the wave (DPP) into wsum[s - 1][wave].
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int s = 1; s <= C; ++s) {
        if (s > ns) break;
        float *B = plane + (s & 1) * PL;
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.active) B[(st.r0 + k) * RX + st.x] = p[k];
        __syncthreads();
        double racc = 0.0;
        if (st.active) {
            const float top = B[(st.r0 > 0 ? st.r0 - 1 : 0) * RX + st.x];
            const float bot = B[(st.r0 + S < RY ? st.r0 + S : RY - 1) * RX + st.x];
            const float *Bl = B + st.r0 * RX + st.xl, *Br = B + st.r0 * RX + st.xr;
            float prev = top;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const float cur = p[k];
                const float below = k == S - 1 ? bot : p[k + 1];
                float sv, kv;
                pressure_sum_comp(Bl[k * RX], Br[k * RX], prev, below, rhs[k], sv, kv);
                p[k] = (sv - kv) * 0.25f;
                const double r = fma(-4.0, (double)cur, (double)sv) - (double)kv;
                const double rm = st.mine(k, a.w, a.h) ? r : 0.0;
                racc = fma(rm, rm, racc);
                prev = cur;
            }
        }
        racc = wave_scan_dpp(racc);
        if (lane == 63) wsum[s - 1][wv] = racc;
    }

    double rK = 0.0;
    if (!fin) {
        float *__restrict__ out = a.pl[(a.launch + 1) % 3];
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.mine(k, a.w, a.h)) out[st.cell(k, a.w)] = p[k];
    } else {
        // Gradient subtraction (src/fluid.cpp:273-295) and the outputs; the
        // plane written here was last read in sweep ns - 1.  The iterate
        // before the last sweep is what this thread published in it (its own
        // slots of plane ns & 1, which nothing has overwritten).
        float pp[S];
        {
            const float *Bl = plane + (ns & 1) * PL;
#pragma unroll
            for (int k = 0; k < S; ++k) pp[k] = st.active ? Bl[(st.r0 + k) * RX + st.x] : 0.0f;
        }
        float *B = plane + ((ns + 1) & 1) * PL;
        if (!kEarlyE) {
#pragma unroll
            for (int k = 0; k < S; ++k) e[k] = a.E[st.mine(k, a.w, a.h) ? st.cell(k, a.w) : 0];
        }
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.active) B[(st.r0 + k) * RX + st.x] = p[k];
        __syncthreads();
        float4 *__restrict__ P4 = reinterpret_cast<float4 *>(a.P);
        float4 *__restrict__ Q4 = reinterpret_cast<float4 *>(a.Q);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (!st.mine(k, a.w, a.h)) continue;
            const int o = (st.r0 + k) * RX + st.x;
            const float pL = B[o - 1], pR = B[o + 1], pT = B[o - RX], pB = B[o + RX];
            const float gradX = (pR - pL) * a.dt / 2.0f;
            const float gradY = (pB - pT) * a.dt / 2.0f;
            const int cell = st.cell(k, a.w);
            P4[cell] = make_float4(e[k].x - gradX, e[k].y - gradY, pp[k], rhs[k]);
            Q4[cell] = make_float4(e[k].x, e[k].y, p[k], rhs[k]);
            // Residual of p^K (the solve ran every sweep): FP64 from the floats.
            const double r = fma(-4.0, (double)p[k],
                                 (double)rhs[k] + (((double)pL + (double)pR) +
                                                   ((double)pT + (double)pB)));
            rK = fma(r, r, rK);
        }
    }

    // ---- this workgroup's FP64 sums into its slots -------------------------
    if (!finish) {
        const double tk = wave_sum(rK), tb = wave_sum(b2);
        if (lane == 0) {
            wsum[C][wv] = tk;
            wsum[C + 1][wv] = tb;
        }
        __syncthreads();
        if ((int)threadIdx.x < NSLOT) {
            double t = 0.0;
            for (int w = 0; w < NW; ++w) t += wsum[threadIdx.x][w];
            rowsum[threadIdx.x] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int m = 0; m < ns; ++m) a.part[(size_t)(b0 + m) * a.ntiles + tile] = rowsum[m];
            if (fin) a.part[(size_t)a.K * a.ntiles + tile] = rowsum[C];
            if (a.launch == 0) a.part[(size_t)kRowB2 * a.ntiles + tile] = rowsum[C + 1];
        }
    }
    // ---- the last workgroup of the launch folds and decides -----------------
    if (threadIdx.x == 0) {
        __threadfence();  // release: this workgroup's slots before its count
        const unsigned n = atomicAdd(&a.ctl->count[a.launch], 1u);
        last_group = n == (unsigned)a.ntiles - 1u;
    }
    __syncthreads();
    if (!last_group) return;
    __threadfence();  // acquire: every workgroup's slots
    if (finish) {
        if (threadIdx.x == 0) {
            a.ctl->kp = stop;
            a.ctl->done = 1;
        }
        return;
    }
    // Slot m of this launch is row b0 + m (m < ns), then row K (fin), then
    // b^2 (launch 0).  Wave w folds slots w, w + NW, ... over the tiles:
    // lane j sums tiles j, j + 64, ... with four accumulators in a fixed
    // pattern, then the butterfly; every total is the same on every run.
    const int nrow = ns + (fin ? 1 : 0) + (a.launch == 0 ? 1 : 0);
    for (int m = wv; m < nrow; m += NW) {
        const int row = m < ns ? b0 + m : (fin && m == ns ? a.K : kRowB2);
        const double *src = a.part + (size_t)row * a.ntiles;
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        int i = lane;
        for (; i + 192 < a.ntiles; i += 256) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                t[q] += __hip_atomic_load(src + i + 64 * q, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int q = 0; i < a.ntiles; i += 64, ++q)
            t[q] += __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double tot = wave_sum((t[0] + t[1]) + (t[2] + t[3]));
        if (lane == 0) {
            a.ctl->total[row] = tot;
            rowsum[m] = tot;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b2 = a.launch == 0 ? rowsum[nrow - 1] : a.ctl->total[kRowB2];
        const double lim = a.tol2 * b2;
        int found = -1;
        for (int m = b0 >= 1 ? 0 : 1; m < ns; ++m)
            if (rowsum[m] <= lim) {
                found = b0 + m;
                break;
            }
        if (found >= 0) {
            a.ctl->stop = found;
        } else if (fin) {
            a.ctl->kp = a.K;
            a.ctl->done = 1;
        }
    }
}

// ---- host dispatch --------------------------------------------------------------
using DiffuseFn = void (*)(DiffuseArgs, PCtl *);
using PressureFn = void (*)(PreciseArgs);

// Diffusion launches: chunks of C = 4 or 8 sweeps (halo C), the last one
// of 1..C sweeps (halo 2..C+1).
template <int TS>
DiffuseFn diffuse_precise_ts(int H, bool last) {
    if (!last) return H == 4 ? k_diffuse_precise<4, false, TS>
                             : H == 8 ? k_diffuse_precise<8, false, TS> : nullptr;
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return k_diffuse_precise<h, true, TS>;
        PFS_CASE(2) PFS_CASE(3) PFS_CASE(4) PFS_CASE(5) PFS_CASE(6) PFS_CASE(7) PFS_CASE(8)
        PFS_CASE(9)
#undef PFS_CASE
    }
    return nullptr;
}

DiffuseFn diffuse_precise_for(int H, bool last, int ts) {
    return ts == 16 ? diffuse_precise_ts<16>(H, last)
                    : ts == 32 ? diffuse_precise_ts<32>(H, last) : diffuse_precise_ts<64>(H, last);
}

template <int H>
PressureFn pressure_precise_h(int ts, bool half) {
    if (ts == 64 && half) return k_pressure_precise<H, 64, 512>;
    return ts == 16 ? k_pressure_precise<H, 16, 1024>
                    : ts == 32 ? k_pressure_precise<H, 32, 1024> : k_pressure_precise<H, 64, 1024>;
}

PressureFn pressure_precise_for(int C, int ts, bool half) {
    return C == 4 ? pressure_precise_h<5>(ts, half) : pressure_precise_h<9>(ts, half);
}

template <int TS>
int threads_precise_ts(int H) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return Geo<h, TS>::NT;
        PFS_CASE(2) PFS_CASE(3) PFS_CASE(4) PFS_CASE(5) PFS_CASE(6) PFS_CASE(7) PFS_CASE(8)
        PFS_CASE(9)
#undef PFS_CASE
    }
    return 0;
}

int threads_precise(int H, int ts) {
    return ts == 16 ? threads_precise_ts<16>(H)
                    : ts == 32 ? threads_precise_ts<32>(H) : threads_precise_ts<64>(H);
}

int threads_pressure_precise(int C, int ts, bool half) {
    if (ts == 64 && half) return C == 4 ? Geo<5, 64, 512>::NT : Geo<9, 64, 512>::NT;
    return threads_precise(C + 1, ts);
}

// Sweeps per launch (4 or 8) and tile side (16, 32, 64; 0: the plain
// step's rule), process-wide.
std::atomic<int> g_precise_chunk{8};
std::atomic<int> g_precise_tile{0};
// 1: 512-thread pressure workgroups at tile side 64 (two per CU: one's
// loads and barriers overlap the other's sweeps; 0.37 against 0.42 ms per
// 2048^2 step with 1024).
std::atomic<int> g_precise_half{1};

// Workspace: the plain fused step's 8 planes (pressure start, b, two float2
// diffusion planes, two pressure planes), the control block, and the
// per-tile sums for the most tiles any side gives.
constexpr size_t kPCtlBytes = (sizeof(PCtl) + 255) / 256 * 256;
size_t precise_ws_bytes(int w, int h) {
    const size_t N = (size_t)w * h;
    const size_t planes = ((8 * N * sizeof(float)) + 255) / 256 * 256;
    return planes + kPCtlBytes + (size_t)kNRows * (size_t)tile_count(w, h, 16) * sizeof(double) +
           256;
}

int precise_step(float *P, float *Q, int w, int h, float dt, float visc, int K, double tol,
                 void *ws, hipStream_t s, PCtl **ctl_out) {
    const int C = g_precise_chunk.load();
    const int forced = g_precise_tile.load();
    const int ts = forced ? forced : tile_for_counts(tile_count(w, h, 64), tile_count(w, h, 32));
    const size_t N = (size_t)w * h;
    char *base = reinterpret_cast<char *>(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    float *wsf = reinterpret_cast<float *>(base);
    PCtl *ctl = reinterpret_cast<PCtl *>(base + ((8 * N * sizeof(float)) + 255) / 256 * 256);
    double *part = reinterpret_cast<double *>(reinterpret_cast<char *>(ctl) + kPCtlBytes);
    *ctl_out = ctl;

    DiffuseArgs d = diffuse_args(P, Q, w, h, dt, visc, K, wsf, ts);
    Chunk plan[kMaxK];
    const int nd = chunk_plan(K, plan, kMaxK, C);
    for (int c = 0; c < nd; ++c) {
        diffuse_chunk(d, c, plan[c], wsf);
        hipLaunchKernelGGL(diffuse_precise_for(plan[c].H, plan[c].last, ts), dim3(d.ntiles),
                           dim3(threads_precise(plan[c].H, ts)), 0, s, d, c == 0 ? ctl : nullptr);
        PFS_TRY(check_launch("precise diffuse"));
    }
    PreciseArgs pa{};
    pa.P = P;
    pa.Q = Q;
    pa.pl[0] = wsf;  // d.pw: the pressure start
    pa.pl[1] = wsf + 6 * N;
    pa.pl[2] = wsf + 7 * N;
    pa.bpl = wsf + N;
    pa.E = reinterpret_cast<const float2 *>(wsf + 2 * N + 2 * N * ((nd - 1) & 1));
    pa.part = part;
    pa.ctl = ctl;
    pa.w = w;
    pa.h = h;
    pa.tiles_x = d.tiles_x;
    pa.ntiles = d.ntiles;
    pa.dt = dt;
    pa.K = K;
    pa.C = C;
    pa.nchunks = (K + C - 1) / C;
    pa.tol2 = tol * tol;
    const bool half = g_precise_half.load() != 0;
    const PressureFn pf = pressure_precise_for(C, ts, half);
    const int nt = threads_pressure_precise(C, ts, half);
    for (int l = 0; l <= pa.nchunks; ++l) {  // the chunks, then the fix-up
        pa.launch = l;
        hipLaunchKernelGGL(pf, dim3(pa.ntiles), dim3(nt), 0, s, pa);
        PFS_TRY(check_launch("precise pressure"));
    }
    return PFS_OK;
}

}  // namespace
}  // namespace pfs

using namespace pfs;

extern "C" {

size_t pfs_precise_workspace_bytes(int vx, int vy, int fp64) {
    if (vx < 1 || vy < 1) return 0;
    return fp64 ? twin_ws_bytes(vx, vy) : precise_ws_bytes(vx, vy);
}

int pfs_set_precise_schedule(int sweeps_per_launch, int tile, int pressure_threads) {
    if (sweeps_per_launch != 0 && sweeps_per_launch != 4 && sweeps_per_launch != 8)
        return set_error(PFS_ERR_INVALID, "sweeps per launch must be 0 (default), 4 or 8");
    if (tile != 0 && !valid_tile_side(tile))
        return set_error(PFS_ERR_INVALID, "tile side must be 0 (automatic), 16, 32 or 64");
    if (pressure_threads != 0 && pressure_threads != 512 && pressure_threads != 1024)
        return set_error(PFS_ERR_INVALID, "pressure threads must be 0 (default), 512 or 1024");
    g_precise_chunk = sweeps_per_launch ? sweeps_per_launch : 8;
    g_precise_tile = tile;
    g_precise_half = pressure_threads != 1024 ? 1 : 0;
    return PFS_OK;
}

int pfs_simulate_fluid_step_precise(pfs_field *vp, pfs_field *tmp, float dt, float viscosity,
                                    int max_iters, double tol, void *workspace,
                                    size_t workspace_bytes, void *stream,
                                    pfs_precise_report *report) {
    PFS_TRY(check_field(vp, "vp", 4));
    PFS_TRY(check_field(tmp, "tmp", 4));
    if (vp->x != tmp->x || vp->y != tmp->y || vp->z != 4 || tmp->z != 4)
        return set_error(PFS_ERR_INVALID, "vp and tmp must both be W x H x 4");
    if (vp->data == tmp->data) return set_error(PFS_ERR_INVALID, "vp and tmp alias");
    if ((uintptr_t)vp->data % 16 || (uintptr_t)tmp->data % 16)
        return set_error(PFS_ERR_UNSUPPORTED, "vp and tmp must be 16-byte aligned");
    if (max_iters < 1 || max_iters > kMaxK)
        return set_error(PFS_ERR_INVALID, "max_iters must be in [1, PFS_PRECISE_MAX_ITERS]");
    if (!(tol >= 0.0)) return set_error(PFS_ERR_INVALID, "tol must be >= 0");
    if (!workspace || workspace_bytes < precise_ws_bytes(vp->x, vp->y))
        return set_error(PFS_ERR_INVALID, "workspace smaller than pfs_precise_workspace_bytes()");
    hipStream_t s = (hipStream_t)stream;
    PCtl *ctl = nullptr;
    PFS_TRY(precise_step(vp->data, tmp->data, vp->x, vp->y, dt, viscosity, max_iters, tol,
                         workspace, s, &ctl));
    if (report) {
        PCtl h;
        PFS_HIP_TRY(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, s));
        PFS_HIP_TRY(hipStreamSynchronize(s));
        std::memset(report, 0, sizeof(*report));
        report->iters_used = h.kp;
        report->b_norm = std::sqrt(h.total[kRowB2]);
        report->residual_norm = std::sqrt(h.total[h.kp]);
        for (int k = 0; k <= kMaxK; ++k) report->residual_history[k] = std::sqrt(h.total[k]);
    }
    return PFS_OK;
}

int pfs_simulate_fluid_step_f64(double *vp, double *tmp, int vx, int vy, double dt,
                                double viscosity, int iters_diffuse, int iters_pressure,
                                void *workspace, size_t workspace_bytes, void *stream) {
    if (!vp || !tmp || vp == tmp) return set_error(PFS_ERR_INVALID, "vp/tmp NULL or aliased");
    if (vx < 1 || vy < 1 || (long long)vx * vy * 4 > (long long)INT32_MAX)
        return set_error(PFS_ERR_INVALID, "bad field size");
    if (iters_diffuse < 1 || iters_pressure < 1 || iters_pressure > iters_diffuse ||
        iters_diffuse > kMaxK)
        return set_error(PFS_ERR_INVALID, "need 1 <= iters_pressure <= iters_diffuse <= max");
    if (!workspace || workspace_bytes < twin_ws_bytes(vx, vy))
        return set_error(PFS_ERR_INVALID, "workspace smaller than pfs_precise_workspace_bytes()");
    return twin_step(vp, tmp, vx, vy, dt, viscosity, iters_diffuse, iters_pressure, workspace,
                     (hipStream_t)stream);
}

}  // extern "C"
