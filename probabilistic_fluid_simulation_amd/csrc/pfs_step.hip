// Fused, temporally blocked fluid step for gfx950.
//
// One call of the reference's simulate_fluid_step (src/fluid.cpp:298-310)
// is 2K+3 full-grid passes there.  Here it is two kernels (for K <= 8):
//
//   D: advect (src/fluid.cpp:24-70) + K diffusion sweeps (:129-196) +
//      divergence (:221-237), temporally blocked in LDS;
//   P: K pressure sweeps (:239-266) + gradient subtraction (:269-296),
//      temporally blocked in LDS, optionally followed by the dye advection
//      of the same cells (:72-127) when one call runs the whole timestep of
//      the reference driver (src/main.cpp:222-225).
//
// Each workgroup owns a TX x TY tile of cells (64, 32 or 16 square: the
// caller's choice) and stages the tile plus a
// halo of H = sweeps (+1 for the final stencil) cells, periodic wrap folded
// into the staged coordinates.  Thread (x, seg) owns column x of the staged
// region, rows [seg*S, seg*S + S), in registers.  Every sweep it publishes
// its cells (alpha-scaled for diffusion) to an LDS plane, waits at one
// barrier (planes alternate), and updates its cells from the left/right
// columns in LDS and its own registers above and below.  Cells of the outer
// ring read clamped neighbours and become garbage; that ring grows by one
// cell per sweep exactly as the valid region shrinks, so the tile is never
// touched by it.  Halo cells are recomputed by both neighbouring tiles with
// the same operations in the same order: every value is bit-identical to the
// reference's global sweep.  For K > 8 the sweeps run in chunks of 8 with the
// iterate in workspace planes between chunks.
//
// Buffer contract (equal to fluid.cpp's, derived from its pointer swaps):
// with P = vp and Q = tmp at entry, E = D^K if K is even else D^(K-1),
// warm = Q[2] if K is even else P[2] (the pressure start), p^0 = warm:
//   Q = (E.u, E.v, p^K, b)     P = (E - grad p^K, p^(K-1), b)
// b = divergence of D^K.  Kernel D only reads vp and tmp: it writes the
// pressure start, b and E to workspace planes (16 B per cell), so that
// kernel P's halo reads never race with its own writes, and kernel P writes
// both outputs whole (32 B per cell) without reading them.

#include <cmath>

#include "pfs_internal.h"
#include "pfs_step_kernels.h"

namespace pfs {

namespace {

template <int H, bool LAST, int TS>
__global__ __launch_bounds__((Geo<H, TS>::NT)) void k_diffuse(DiffuseArgs a) {
    diffuse_tile<H, LAST, TS>(a, xcd_tile(blockIdx.x, a.ntiles));
}

template <int H, bool LAST, int TS>
__global__ __launch_bounds__((Geo<H, TS>::NT)) void k_pressure(PressureArgs a) {
    pressure_tile<H, LAST, TS>(a, xcd_tile(blockIdx.x, a.ntiles));
}

// ---- host dispatch over the compile-time halo and tile side ----------------
using DiffuseFn = void (*)(DiffuseArgs);
using PressureFn = void (*)(PressureArgs);

#define PFS_H_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9)

template <int TS>
DiffuseFn diffuse_for_ts(int H, bool last) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return last ? k_diffuse<h, true, TS> : k_diffuse<h, false, TS>;
        PFS_H_CASES(PFS_CASE)
#undef PFS_CASE
    }
    return nullptr;
}

template <int TS>
PressureFn pressure_for_ts(int H, bool last) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return last ? k_pressure<h, true, TS> : k_pressure<h, false, TS>;
        PFS_H_CASES(PFS_CASE)
#undef PFS_CASE
    }
    return nullptr;
}

DiffuseFn diffuse_for(int H, bool last, int ts) {
    return ts == 16 ? diffuse_for_ts<16>(H, last)
                    : ts == 32 ? diffuse_for_ts<32>(H, last) : diffuse_for_ts<64>(H, last);
}

PressureFn pressure_for(int H, bool last, int ts) {
    return ts == 16 ? pressure_for_ts<16>(H, last)
                    : ts == 32 ? pressure_for_ts<32>(H, last) : pressure_for_ts<64>(H, last);
}

template <int TS>
int threads_for_ts(int H) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return Geo<h, TS>::NT;
        PFS_H_CASES(PFS_CASE)
#undef PFS_CASE
    }
    return 0;
}

}  // namespace

int threads_for(int H, int ts) {
    return ts == 16 ? threads_for_ts<16>(H) : ts == 32 ? threads_for_ts<32>(H) : threads_for_ts<64>(H);
}

size_t fused_workspace_bytes(int w, int h) {
    return scene_ws_floats(w, h, kMaxChunk + 1) * sizeof(float);
}

int fused_step(const StepArgs &s, float *ws, hipStream_t stream, hipEvent_t *ev, int reps) {
    if (s.iters < 1) return set_error(PFS_ERR_INVALID, "fused step needs iters >= 1");
    Chunk plan[64];
    const int nchunks = chunk_plan(s.iters, plan, 64);
    if (plan[nchunks - 1].last == false)
        return set_error(PFS_ERR_UNSUPPORTED, "iters too large for one step plan");
    if (!valid_tile_side(s.tile)) return set_error(PFS_ERR_INVALID, "tile side must be 16, 32 or 64");
    DiffuseArgs d = diffuse_args(s.P, s.Q, s.w, s.h, s.dt, s.visc, s.iters, ws, s.tile);
    PressureArgs p = pressure_args(s.P, s.Q, s.image, s.itmp, s.w, s.h, s.dt, ws, s.tile);
    const dim3 grid(d.ntiles);

    if (!ev) reps = 1;
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[0], stream));
    for (int r = 0; r < reps; ++r)
        for (int c = 0; c < nchunks; ++c) {
            diffuse_chunk(d, c, plan[c], ws);
            hipLaunchKernelGGL(diffuse_for(plan[c].H, plan[c].last, s.tile), grid,
                               dim3(threads_for(plan[c].H, s.tile)), 0, stream, d);
            PFS_TRY(check_launch("fused diffuse"));
        }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[1], stream));
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[2], stream));
    for (int r = 0; r < reps; ++r)
        for (int c = 0; c < nchunks; ++c) {
            pressure_chunk(p, c, plan[c], ws, s.image, s.itmp);
            hipLaunchKernelGGL(pressure_for(plan[c].H, plan[c].last, s.tile), grid,
                               dim3(threads_for(plan[c].H, s.tile)), 0, stream, p);
            PFS_TRY(check_launch("fused pressure"));
        }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[3], stream));
    return PFS_OK;
}

}  // namespace pfs
