// Precision form of the fluid step, and its FP64 twin.
//
// No counterpart in the reference, whose Jacobi loops run a fixed count in
// FP32 (src/fluid.cpp:151,239).  The precision form keeps the reference's
// step structure (src/fluid.cpp:298-310) and changes three things:
//
//   - the 5-term stencil sums of both Jacobi sweeps (src/fluid.cpp:19-21)
//     are compensated: products by TwoProduct (FMA), sums by TwoSum, so the
//     sum is the float nearest its value in about twice FP32 precision
//     (Ogita, Rump & Oishi 2005, Dot2/Sum2) before the single rounding and
//     the division;
//   - every pressure sweep also accumulates, in FP64, the residual of its
//     input iterate, r = b + pL + pR + pT + pB - 4p (zero at the fixed point
//     of the sweep p' = (pL+pR+pT+pB+b)/4), and sum b^2 once;
//   - the pressure iteration stops after the first iterate whose residual
//     satisfies ||r|| <= tol ||b||, at most K sweeps.  The test runs on the
//     device: sweep s reads the completed residual of iterate s-2 and skips
//     itself once converged, so no sweep waits on the host.
//
// Buffer contract, with Kd = K the diffusion sweeps and Kp <= K the pressure
// sweeps actually run: E = D^Kd if Kd is even else D^(Kd-1); the pressure
// starts from tmp[2] if Kd is even else vp[2];
//   tmp = (E.u, E.v, p^Kp, b)     vp = (E - grad p^Kp, p^(Kp-1), b)
// which is exactly the reference's step when Kp = Kd (the plain step's
// contract, see pfs_step.hip).
//
// The FP64 twin runs the same contract in double with plain arithmetic and
// given (Kd, Kp): the yardstick the precision form's error is reported
// against.  With Kd = Kp = K it reproduces the reference's algorithm in
// double exactly.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "pfs.h"
#include "pfs_internal.h"

namespace pfs {
namespace {

constexpr int kBX = 64, kBY = 4;
constexpr int kRows = 8;  // rows per thread: a block covers kBX x (kBY * kRows) cells
constexpr int kMaxK = PFS_PRECISE_MAX_ITERS;

template <typename T>
struct V2;
template <>
struct V2<float> {
    using type = float2;
};
template <>
struct V2<double> {
    using type = double2;
};

// FP64 sums are accumulated with one atomic per block, spread over kSlots
// addresses a cache line apart (slot = block index mod kSlots) so that
// blocks do not all queue on one address; a reader adds the slots in
// slot order.
constexpr int kSlots = 32;
constexpr int kSlotStride = 16;     // doubles: 128 bytes
constexpr int kSumB2 = kMaxK + 1;   // row of sum b^2
constexpr int kSumFinal = kMaxK + 2;  // row of sum r^2 of the kept iterate
constexpr int kSums = kMaxK + 3;    // rows 0..kMaxK: sum r^2 of pressure iterate s

// Device-side state of one precision step (cleared by k_ctl_clear).
struct Ctl {
    double sum[kSums][kSlots * kSlotStride];
    int done, kp;
};

// The folded totals the host reads back (written by k_ctl_fold).
struct CtlTotals {
    double sum[kSums];
    int kp, pad;
};

__device__ __forceinline__ double ctl_total(const Ctl *c, int row) {
    double t = 0.0;
#pragma unroll 8
    for (int k = 0; k < kSlots; ++k) t += c->sum[row][k * kSlotStride];
    return t;
}

// ---- compensated arithmetic (error-free transformations) -------------------
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// alpha*(l + r + t + b) + c summed as written in src/fluid.cpp:175-182 but
// compensated: the float nearest the exactly rounded-once value.
__device__ __forceinline__ float diffuse_sum_comp(float l, float r, float t, float b, float c,
                                                  float alpha) {
    float s = alpha * l;
    float err = fmaf(alpha, l, -s);
    const float x[3] = {r, t, b};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float p = alpha * x[k];
        const float ep = fmaf(alpha, x[k], -p);
        float e;
        two_sum(s, p, s, e);
        err += e + ep;
    }
    float e;
    two_sum(s, c, s, e);
    err += e;
    return s + err;
}

__device__ __forceinline__ float pressure_sum_comp(float l, float r, float t, float b, float rhs) {
    float s = l, err = 0.0f, e;
    two_sum(s, r, s, e);
    err += e;
    two_sum(s, t, s, e);
    err += e;
    two_sum(s, b, s, e);
    err += e;
    two_sum(s, rhs, s, e);
    err += e;
    return s + err;
}

// Block-wide FP64 sum into row `row` of the control sums (one atomic per
// block).
__device__ __forceinline__ void block_add(double v, Ctl *ctl, int row) {
    __shared__ double part[kBX * kBY / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int tid = threadIdx.y * kBX + threadIdx.x;
    if ((tid & 63) == 0) part[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int k = 0; k < kBX * kBY / 64; ++k) s += part[k];
        const int slot = (int)((blockIdx.y * gridDim.x + blockIdx.x) % kSlots);
        atomicAdd(&ctl->sum[row][slot * kSlotStride], s);
    }
}

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

// Grid (kSums): block r clears row r; block 0 also the decision.
__global__ void k_ctl_clear(Ctl *c) {
    for (int k = threadIdx.x; k < kSlots * kSlotStride; k += blockDim.x) c->sum[blockIdx.x][k] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->done = 0;
        c->kp = 0;
    }
}

// One block of 128 threads: thread r < kSums folds row r.
static_assert(kSums <= 128, "k_ctl_fold folds one row per thread");
__global__ void k_ctl_fold(const Ctl *c, CtlTotals *t) {
    if ((int)threadIdx.x < kSums) t->sum[threadIdx.x] = ctl_total(c, threadIdx.x);
    if (threadIdx.x == 0) t->kp = c->kp;
}

// Advection of vp[0,1] (src/fluid.cpp:24-70) into the iterate plane d.
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void k_p_advect(const T *__restrict__ P,
                                                      typename V2<T>::type *__restrict__ d, int w,
                                                      int h, T dt, T fw, T fh, T inv_w, T inv_h,
                                                      int pow2_w, int pow2_h) {
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const T xp = (T)c.i - div_n(dt * P[4 * cell], fw, inv_w, pow2_w);
        const T yp = (T)c.j - div_n(dt * P[4 * cell + 1], fh, inv_h, pow2_h);
        const TraceT<T> t = trace(xp, yp, w, h, fw, fh);
        const int a00 = 4 * (t.j0 * w + t.i0), a10 = 4 * (t.j0 * w + t.i1);
        const int a01 = 4 * (t.j1 * w + t.i0), a11 = 4 * (t.j1 * w + t.i1);
        d[cell] = {bilerp(t.sx, t.sy, P[a00], P[a10], P[a01], P[a11]),
                   bilerp(t.sx, t.sy, P[a00 + 1], P[a10 + 1], P[a01 + 1], P[a11 + 1])};
    });
}

// One diffusion sweep (src/fluid.cpp:154-185) on the iterate planes.
template <typename T, bool COMP>
__global__ __launch_bounds__(kBX *kBY) void k_p_diffuse(const typename V2<T>::type *__restrict__ in,
                                                       typename V2<T>::type *__restrict__ out,
                                                       int w, int h, T alpha, T beta) {
    for_cells(w, h, [&](const Cell &c) {
        const auto L = in[c.j * w + c.im], R = in[c.j * w + c.ip], Tp = in[c.jm * w + c.i],
                   B = in[c.jp * w + c.i], C = in[c.j * w + c.i];
        if constexpr (COMP) {
            out[c.j * w + c.i] = {diffuse_sum_comp(L.x, R.x, Tp.x, B.x, C.x, alpha) / beta,
                                  diffuse_sum_comp(L.y, R.y, Tp.y, B.y, C.y, alpha) / beta};
        } else {
            out[c.j * w + c.i] = {diffuse_cell(L.x, R.x, Tp.x, B.x, C.x, alpha, beta),
                                  diffuse_cell(L.y, R.y, Tp.y, B.y, C.y, alpha, beta)};
        }
    });
}

// Divergence of D^Kd (src/fluid.cpp:221-237) into b, and the pressure start
// (tmp[2] for even Kd, vp[2] for odd) into iterate plane 0.
template <typename T, bool COMP>
__global__ __launch_bounds__(kBX *kBY) void k_p_divergence(const typename V2<T>::type *__restrict__ dk,
                                                          const T *__restrict__ P,
                                                          const T *__restrict__ Q, T *__restrict__ b,
                                                          T *__restrict__ p0, int w, int h,
                                                          T gamma, int kd_even, Ctl *ctl) {
    double b2 = 0.0;
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const T uR = dk[c.j * w + c.ip].x - dk[c.j * w + c.im].x;
        const T vT = dk[c.jp * w + c.i].y - dk[c.jm * w + c.i].y;
        const T bb = gamma * (uR + vT);
        b[cell] = bb;
        p0[cell] = kd_even ? Q[4 * cell + 2] : P[4 * cell + 2];
        b2 += (double)bb * (double)bb;
    });
    if constexpr (COMP) block_add(b2, ctl, kSumB2);
}

// Pressure sweep s (src/fluid.cpp:240-257): p^s from p^(s-1).  With COMP,
// also the FP64 residual of p^(s-1) into row s-1, and the convergence test
// of iterate s-2, whose residual the previous sweep completed.
template <typename T, bool COMP>
__global__ __launch_bounds__(kBX *kBY) void k_p_pressure(const T *__restrict__ in,
                                                        T *__restrict__ out,
                                                        const T *__restrict__ b, int w, int h, int s,
                                                        double tol2, Ctl *ctl) {
    if constexpr (COMP) {
        // One thread decides for the block; every block reaches the same
        // verdict from the same completed sums, added in the same order.
        __shared__ int skip;
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const bool converged =
                s >= 3 && ctl_total(ctl, s - 2) <= tol2 * ctl_total(ctl, kSumB2);
            skip = ctl->done || converged;
            if (converged && !ctl->done) {
                ctl->kp = s - 2;
                ctl->done = 1;
            }
        }
        __syncthreads();
        if (skip) return;
    }
    double r2 = 0.0;
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const T pL = in[c.j * w + c.im], pR = in[c.j * w + c.ip];
        const T pT = in[c.jm * w + c.i], pB = in[c.jp * w + c.i];
        const T bb = b[cell];
        if constexpr (COMP) {
            out[cell] = pressure_sum_comp(pL, pR, pT, pB, bb) * 0.25f;
            const double r = (double)bb + (double)pL + (double)pR + (double)pT + (double)pB -
                             4.0 * (double)in[cell];
            r2 += r * r;
        } else {
            out[cell] = pressure_cell(pL, pR, pT, pB, bb);
        }
    });
    if constexpr (COMP) block_add(r2, ctl, s - 1);
}

// Gradient subtraction with p^Kp (src/fluid.cpp:269-296) and the output
// contract.  COMP: Kp comes from the device-side test (with the last
// iterate's residual checked here); otherwise kp_given.
template <typename T, bool COMP>
__global__ __launch_bounds__(kBX *kBY) void k_p_finalize(
    T *__restrict__ P, T *__restrict__ Q, const typename V2<T>::type *__restrict__ dk,
    const typename V2<T>::type *__restrict__ dkm1, const T *p0, const T *p1, const T *p2,
    const T *__restrict__ b, int w, int h, T dt, int kd, int kp_given, double tol2, Ctl *ctl) {
    int kp = kp_given;
    if constexpr (COMP) {
        __shared__ int kp_s;
        if (threadIdx.x == 0 && threadIdx.y == 0)
            kp_s = ctl->done ? ctl->kp
                             : ((kd >= 2 && ctl_total(ctl, kd - 1) <= tol2 * ctl_total(ctl, kSumB2))
                                    ? kd - 1
                                    : kd);
        __syncthreads();
        kp = kp_s;
    }
    // Iterate k lives in plane k % 3.
    const T *pk = kp % 3 == 0 ? p0 : kp % 3 == 1 ? p1 : p2;
    const T *pkm1 = kp % 3 == 1 ? p0 : kp % 3 == 2 ? p1 : p2;
    double r2 = 0.0;
    for_cells(w, h, [&](const Cell &c) {
        const int cell = c.j * w + c.i;
        const T pL = pk[c.j * w + c.im], pR = pk[c.j * w + c.ip];
        const T pT = pk[c.jm * w + c.i], pB = pk[c.jp * w + c.i];
        const T gradX = (pR - pL) * dt / (T)2.0f;
        const T gradY = (pB - pT) * dt / (T)2.0f;
        const auto e = (kd % 2 == 0) ? dk[cell] : dkm1[cell];
        const T bb = b[cell];
        P[4 * cell] = e.x - gradX;
        P[4 * cell + 1] = e.y - gradY;
        P[4 * cell + 2] = pkm1[cell];
        P[4 * cell + 3] = bb;
        Q[4 * cell] = e.x;
        Q[4 * cell + 1] = e.y;
        Q[4 * cell + 2] = pk[cell];
        Q[4 * cell + 3] = bb;
        const double r = (double)bb + (double)pL + (double)pR + (double)pT + (double)pB -
                         4.0 * (double)pk[cell];
        r2 += r * r;
    });
    if constexpr (COMP) {
        block_add(r2, ctl, kSumFinal);
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0)
            ctl->kp = kp;
    }
}

template <typename T, bool COMP>
int precise_step(T *P, T *Q, int w, int h, T dt, T visc, int kd, int kp, double tol, void *ws,
                 hipStream_t s, Ctl *ctl) {
    using T2 = typename V2<T>::type;
    const size_t N = (size_t)w * h;
    T2 *d[2] = {static_cast<T2 *>(ws), static_cast<T2 *>(ws) + N};
    T *b = reinterpret_cast<T *>(d[1] + N);
    T *p[3] = {b + N, b + 2 * N, b + 3 * N};
    const dim3 grid((w + kBX - 1) / kBX, (h + kBY * kRows - 1) / (kBY * kRows)),
        block(kBX, kBY);
    const T alpha = visc * dt;
    const T beta = (T)(1.0 + 4.0 * (double)alpha);  // src/fluid.cpp:145
    const T gamma = (T)(-1.0 / (double)dt);         // src/fluid.cpp:218
    const double tol2 = tol * tol;
    if (COMP) hipLaunchKernelGGL(k_ctl_clear, dim3(kSums), dim3(256), 0, s, ctl);
    hipLaunchKernelGGL((k_p_advect<T>), grid, block, 0, s, P, d[0], w, h, dt, (T)w, (T)h,
                       (T)1 / (T)w, (T)1 / (T)h, (w & (w - 1)) == 0, (h & (h - 1)) == 0);
    for (int k = 1; k <= kd; ++k)
        hipLaunchKernelGGL((k_p_diffuse<T, COMP>), grid, block, 0, s, d[(k - 1) & 1], d[k & 1], w,
                           h, alpha, beta);
    hipLaunchKernelGGL((k_p_divergence<T, COMP>), grid, block, 0, s, d[kd & 1], P, Q, b, p[0], w,
                       h, gamma, kd % 2 == 0, ctl);
    for (int k = 1; k <= kp; ++k)
        hipLaunchKernelGGL((k_p_pressure<T, COMP>), grid, block, 0, s, p[(k - 1) % 3], p[k % 3], b,
                           w, h, k, tol2, ctl);
    hipLaunchKernelGGL((k_p_finalize<T, COMP>), grid, block, 0, s, P, Q, d[kd & 1],
                       d[(kd + 1) & 1], p[0], p[1], p[2], b, w, h, dt, kd, kp, tol2, ctl);
    return check_launch("precision step");
}

// Workspace: two iterate planes of T2, b and three pressure planes of T,
// then (FP32 only) the control block and its folded totals, each 256-byte
// aligned.
constexpr size_t kCtlBytes = (sizeof(Ctl) + 255) / 256 * 256;

template <typename T>
size_t precise_ws_bytes(int w, int h) {
    const size_t N = (size_t)w * h;
    return N * (2 * sizeof(typename V2<T>::type) + 4 * sizeof(T)) + kCtlBytes +
           sizeof(CtlTotals) + 256;
}

}  // namespace
}  // namespace pfs

using namespace pfs;

extern "C" {

size_t pfs_precise_workspace_bytes(int vx, int vy, int fp64) {
    if (vx < 1 || vy < 1) return 0;
    return fp64 ? precise_ws_bytes<double>(vx, vy) : precise_ws_bytes<float>(vx, vy);
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
    if (max_iters < 1 || max_iters > kMaxK)
        return set_error(PFS_ERR_INVALID, "max_iters must be in [1, PFS_PRECISE_MAX_ITERS]");
    if (!workspace || workspace_bytes < precise_ws_bytes<float>(vp->x, vp->y))
        return set_error(PFS_ERR_INVALID, "workspace smaller than pfs_precise_workspace_bytes()");
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)vp->x * vp->y;
    char *ctl_at = static_cast<char *>(workspace) + N * (2 * sizeof(float2) + 4 * sizeof(float));
    Ctl *ctl = reinterpret_cast<Ctl *>(((uintptr_t)ctl_at + 255) & ~(uintptr_t)255);
    PFS_TRY((precise_step<float, true>(vp->data, tmp->data, vp->x, vp->y, dt, viscosity, max_iters,
                                       max_iters, tol, workspace, s, ctl)));
    if (report) {
        CtlTotals *tot = reinterpret_cast<CtlTotals *>(reinterpret_cast<char *>(ctl) + kCtlBytes);
        hipLaunchKernelGGL(k_ctl_fold, dim3(1), dim3(128), 0, s, ctl, tot);
        PFS_TRY(check_launch("precision report"));
        CtlTotals h;
        PFS_HIP_TRY(hipMemcpyAsync(&h, tot, sizeof(h), hipMemcpyDeviceToHost, s));
        PFS_HIP_TRY(hipStreamSynchronize(s));
        std::memset(report, 0, sizeof(*report));
        report->iters_used = h.kp;
        report->b_norm = std::sqrt(h.sum[kSumB2]);
        report->residual_norm = std::sqrt(h.sum[kSumFinal]);
        for (int k = 0; k <= kMaxK; ++k) report->residual_history[k] = std::sqrt(h.sum[k]);
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
    if (!workspace || workspace_bytes < precise_ws_bytes<double>(vx, vy))
        return set_error(PFS_ERR_INVALID, "workspace smaller than pfs_precise_workspace_bytes()");
    return precise_step<double, false>(vp, tmp, vx, vy, dt, viscosity, iters_diffuse,
                                       iters_pressure, 0.0, workspace, (hipStream_t)stream,
                                       nullptr);
}

}  // extern "C"
