// The extern "C" ABI of libpfs.so (include/pfs.h): argument validation,
// the reference's buffer ping-pong for the per-operator entry points, the
// step-path choice and the fused step's device workspace.

#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>

#include "pfs.h"
#include "pfs_internal.h"
#include "pfs_step_kernels.h"

namespace pfs {

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_step_path{PFS_PATH_AUTO};  // process-wide settings, read by every call

// One growable workspace per (device, stream): calls on one stream are
// ordered, so they can share it; calls on different streams must not.
struct Workspace {
    float *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::map<std::pair<int, hipStream_t>, Workspace> g_ws;

int get_workspace(size_t bytes, hipStream_t s, float **out) {
    int dev = 0;
    PFS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Workspace &w = g_ws[{dev, s}];
    if (w.bytes < bytes) {
        if (w.ptr) {
            // Work already queued on this stream may still use the old block.
            PFS_HIP_TRY(hipStreamSynchronize(s));
            PFS_HIP_TRY(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(PFS_ERR_NOMEM, "workspace allocation of " + std::to_string(bytes) +
                                                " bytes failed");
        }
        w.ptr = static_cast<float *>(p);
        w.bytes = bytes;
    }
    *out = w.ptr;
    return PFS_OK;
}

bool same_shape(const pfs_field *a, const pfs_field *b) {
    return a->x == b->x && a->y == b->y && a->z == b->z;
}

// Tile side of the fused step: the side pfs_set_step_tile forced, else the
// side pfs_autotune_step measured fastest for this shape, else the rule.
std::atomic<int> g_step_tile{0};  // 0: not forced
std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int, int>, int> g_tuned;  // (w, h, iters, dye fused) -> side

}  // namespace

int device_cus() {
    static std::mutex m;
    static std::map<int, int> cus;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(m);
    auto it = cus.find(dev);
    if (it != cus.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
        n = 256;
    cus[dev] = n;
    return n;
}

// The rule: the largest side with at least one tile for every two CUs; a
// field too small for that takes the smallest side, whose workgroups each
// hold the fewest rows per thread.  (Measured on MI355X, 256 CUs: 64-wide
// tiles win from 144 tiles up (768^2), 32-wide ones at 512^2, 16-wide ones
// at 256^2 and below.)
int tile_for_counts(long long tiles64, long long tiles32) {
    const long long half = device_cus() / 2;
    return tiles64 >= half ? 64 : tiles32 >= half ? 32 : 16;
}

long long tile_count(int w, int h, int ts) {
    return (long long)((w + ts - 1) / ts) * ((h + ts - 1) / ts);
}

namespace {
int rule_tile(int w, int h) { return tile_for_counts(tile_count(w, h, 64), tile_count(w, h, 32)); }

int step_tile(int w, int h, int iters, bool dye_fused) {
    if (const int forced = g_step_tile.load()) return forced;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find({w, h, iters, dye_fused ? 1 : 0});
        if (it != g_tuned.end()) return it->second;
    }
    return rule_tile(w, h);
}
}  // namespace

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(PFS_ERR_HIP, std::string(what) + " launch: " + hipGetErrorString(e));
    return PFS_OK;
}

int check_field(const pfs_field *f, const char *name, int min_channels) {
    if (!f) return set_error(PFS_ERR_INVALID, std::string(name) + " is NULL");
    if (!f->data) return set_error(PFS_ERR_INVALID, std::string(name) + "->data is NULL");
    if (f->x < 1 || f->y < 1 || f->z < min_channels)
        return set_error(PFS_ERR_INVALID, std::string(name) + " has shape " +
                                              std::to_string(f->x) + "x" + std::to_string(f->y) +
                                              "x" + std::to_string(f->z) + " (needs >= 1x1x" +
                                              std::to_string(min_channels) + ")");
    if ((long long)f->x * f->y * f->z > (long long)INT_MAX)
        return set_error(PFS_ERR_UNSUPPORTED,
                         std::string(name) + " holds 2^31 or more floats (32-bit indexing)");
    return PFS_OK;
}

}  // namespace pfs

using namespace pfs;

extern "C" {

const char *pfs_version(void) { return "pfs 0.1.0 (gfx950)"; }

const char *pfs_last_error(void) { return g_last_error.c_str(); }

int pfs_advect(const pfs_field *vp, pfs_field *vp_out, float dt, void *stream) {
    PFS_TRY(check_field(vp, "vp", 2));
    PFS_TRY(check_field(vp_out, "vp_out", 2));
    if (!same_shape(vp, vp_out)) return set_error(PFS_ERR_INVALID, "vp_out shape != vp shape");
    if (vp->data == vp_out->data) return set_error(PFS_ERR_INVALID, "vp and vp_out alias");
    return launch_advect(vp->data, vp_out->data, vp->x, vp->y, vp->z, dt, (hipStream_t)stream);
}

int pfs_advect_color(const pfs_field *image, pfs_field *out, const pfs_field *vp, float dt,
                     void *stream) {
    PFS_TRY(check_field(image, "image", 1));
    PFS_TRY(check_field(out, "out", 1));
    PFS_TRY(check_field(vp, "vp", 2));
    if (!same_shape(image, out)) return set_error(PFS_ERR_INVALID, "out shape != image shape");
    if (image->data == out->data) return set_error(PFS_ERR_INVALID, "image and out alias");
    if (out->data == vp->data) return set_error(PFS_ERR_INVALID, "out and vp alias");
    return launch_advect_color(image->data, out->data, vp->data, image->x, image->y, image->z,
                               vp->x, vp->y, vp->z, dt, (hipStream_t)stream);
}

int pfs_diffuse(pfs_field *vp, pfs_field *vp_out, float viscosity, float dt, int iters,
                void *stream) {
    PFS_TRY(check_field(vp, "vp", 2));
    PFS_TRY(check_field(vp_out, "vp_out", 2));
    if (!same_shape(vp, vp_out)) return set_error(PFS_ERR_INVALID, "vp_out shape != vp shape");
    if (vp->data == vp_out->data) return set_error(PFS_ERR_INVALID, "vp and vp_out alias");
    if (iters < 0) return set_error(PFS_ERR_INVALID, "iters < 0");
    const float alpha = diffuse_alpha(viscosity, dt);
    const float beta = diffuse_beta(alpha);
    for (int it = 0; it < iters; ++it) {
        PFS_TRY(launch_diffuse_sweep(vp->data, vp_out->data, vp->x, vp->y, vp->z, alpha, beta,
                                     (hipStream_t)stream));
        if (it != iters - 1) std::swap(vp->data, vp_out->data);  // src/fluid.cpp:188-194
    }
    return PFS_OK;
}

int pfs_add_forces(pfs_field *vp, const float *forces, void *stream) {
    (void)forces;
    (void)stream;
    PFS_TRY(check_field(vp, "vp", 1));
    return PFS_OK;  // src/fluid.cpp:198-208 is an empty loop
}

int pfs_add_forces_apply(pfs_field *vp, const float *forces, int fz, float dt, void *stream) {
    PFS_TRY(check_field(vp, "vp", 1));
    if (!forces) return set_error(PFS_ERR_INVALID, "forces is NULL");
    if (fz < 1 || fz > vp->z) return set_error(PFS_ERR_INVALID, "need 1 <= fz <= vp->z");
    if (forces == vp->data) return set_error(PFS_ERR_INVALID, "forces and vp alias");
    return launch_add_forces(vp->data, forces, vp->x, vp->y, vp->z, fz, dt, (hipStream_t)stream);
}

int pfs_compute_pressure(pfs_field *vp, pfs_field *vp_out, float dt, int iters, void *stream) {
    PFS_TRY(check_field(vp, "vp", 4));
    PFS_TRY(check_field(vp_out, "vp_out", 4));
    if (!same_shape(vp, vp_out)) return set_error(PFS_ERR_INVALID, "vp_out shape != vp shape");
    if (vp->data == vp_out->data) return set_error(PFS_ERR_INVALID, "vp and vp_out alias");
    if (iters < 0) return set_error(PFS_ERR_INVALID, "iters < 0");
    hipStream_t s = (hipStream_t)stream;
    PFS_TRY(launch_divergence(vp->data, vp_out->data, vp->x, vp->y, vp->z, pressure_gamma(dt), s));
    for (int it = 0; it < iters; ++it) {
        PFS_TRY(launch_pressure_sweep(vp->data, vp_out->data, vp->x, vp->y, vp->z, s));
        if (it != iters - 1) std::swap(vp->data, vp_out->data);  // src/fluid.cpp:260-265
    }
    return PFS_OK;
}

int pfs_subtract_pressure_gradient(const pfs_field *vp, pfs_field *vp_out, float dt,
                                   void *stream) {
    PFS_TRY(check_field(vp, "vp", 3));
    PFS_TRY(check_field(vp_out, "vp_out", 2));
    if (vp->x != vp_out->x || vp->y != vp_out->y || vp->z != vp_out->z)
        return set_error(PFS_ERR_INVALID, "vp_out shape != vp shape");
    if (vp->data == vp_out->data) return set_error(PFS_ERR_INVALID, "vp and vp_out alias");
    return launch_subtract_gradient(vp->data, vp_out->data, vp->x, vp->y, vp->z, dt,
                                    (hipStream_t)stream);
}

static int step_per_op(pfs_field *vp, pfs_field *tmp, float dt, float viscosity, int iters,
                       void *stream) {
    // src/fluid.cpp:298-310, struct swaps included.
    PFS_TRY(pfs_advect(vp, tmp, dt, stream));
    PFS_TRY(pfs_diffuse(tmp, vp, viscosity, dt, iters, stream));
    PFS_TRY(pfs_compute_pressure(vp, tmp, dt, iters, stream));
    return pfs_subtract_pressure_gradient(tmp, vp, dt, stream);
}

static bool aligned16(const void *p) { return (uintptr_t)p % 16 == 0; }

static int check_step_fields(const pfs_field *vp, const pfs_field *tmp, int iters) {
    PFS_TRY(check_field(vp, "vp", 4));
    PFS_TRY(check_field(tmp, "tmp", 4));
    if (!same_shape(vp, tmp)) return set_error(PFS_ERR_INVALID, "tmp shape != vp shape");
    if (vp->data == tmp->data) return set_error(PFS_ERR_INVALID, "vp and tmp alias");
    if (iters < 0) return set_error(PFS_ERR_INVALID, "iters < 0");
    return PFS_OK;
}

// Sweeps one fused step plans for: 64 launches of kMaxChunk sweeps.
constexpr int kFusedMaxIters = 64 * kMaxChunk;

// Whether the fused kernels apply: 4-channel, 16-byte aligned fields,
// 1 <= K <= kFusedMaxIters (a larger K takes the per-operator sequence,
// which accepts any count, as the reference does).
static bool fused_ok(const pfs_field *vp, const pfs_field *tmp, int iters) {
    return g_step_path == PFS_PATH_AUTO && vp->z == 4 && iters >= 1 && iters <= kFusedMaxIters &&
           aligned16(vp->data) && aligned16(tmp->data);
}

// simulate_fluid_step, optionally followed by advect_color_step on
// (image, itmp) as one timestep of src/main.cpp:222-225.  `ev` (4 events)
// requests per-kernel timing of the fused path.
//
// `ws_given` (with `ws_cap` bytes) replaces the stream's library-allocated
// workspace; a fused step that needs more is rejected before any launch.
static int timestep(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                    float viscosity, int iters, void *stream, hipEvent_t *ev, int reps = 1,
                    int tile = 0, float *ws_given = nullptr, size_t ws_cap = 0) {
    PFS_TRY(check_step_fields(vp, tmp, iters));
    if (image) {
        PFS_TRY(check_field(image, "image", 1));
        PFS_TRY(check_field(itmp, "itmp", 1));
        if (!same_shape(image, itmp)) return set_error(PFS_ERR_INVALID, "itmp shape != image shape");
        if (image->data == itmp->data) return set_error(PFS_ERR_INVALID, "image and itmp alias");
        for (const pfs_field *f : {vp, tmp})
            if (f->data == image->data || f->data == itmp->data)
                return set_error(PFS_ERR_INVALID, "dye buffers alias the velocity buffers");
    }
    if (!fused_ok(vp, tmp, iters)) {
        if (ev) return set_error(PFS_ERR_UNSUPPORTED, "profiling needs the fused path");
        PFS_TRY(step_per_op(vp, tmp, dt, viscosity, iters, stream));
        return image ? pfs_advect_color_step(image, itmp, vp, dt, stream) : PFS_OK;
    }
    // The dye rides in the pressure kernel when pixel (i, j) samples cell
    // (i, j): same size, 4 channels, float4-aligned.
    const bool dye_fused = image && image->x == vp->x && image->y == vp->y && image->z == 4 &&
                           aligned16(image->data) && aligned16(itmp->data);
    float *ws = ws_given;
    if (ws_given) {
        const size_t need = scene_ws_floats(vp->x, vp->y, iters) * sizeof(float);
        if (need > ws_cap)
            return set_error(PFS_ERR_UNSUPPORTED,
                             "a " + std::to_string(vp->x) + "x" + std::to_string(vp->y) +
                                 " field with iters = " + std::to_string(iters) + " needs " +
                                 std::to_string(need) + " bytes of step scratch; the " +
                                 "no-allocation entry point holds " + std::to_string(ws_cap));
    } else {
        PFS_TRY(get_workspace(fused_workspace_bytes(vp->x, vp->y), (hipStream_t)stream, &ws));
    }
    StepArgs a{vp->data, tmp->data, dye_fused ? image->data : nullptr,
               dye_fused ? itmp->data : nullptr, vp->x, vp->y, dt, viscosity, iters,
               tile ? tile : step_tile(vp->x, vp->y, iters, dye_fused)};
    PFS_TRY(fused_step(a, ws, (hipStream_t)stream, ev, reps));
    if (!image) return PFS_OK;
    if (!dye_fused) return pfs_advect_color_step(image, itmp, vp, dt, stream);
    std::swap(image->data, itmp->data);  // src/fluid.cpp:317-319
    return PFS_OK;
}

// An even number of consecutive fused timesteps captured in a hipGraph, kept
// for later calls with the same buffers, shapes, coefficients, tile side and
// workspace.  (Even: the dye buffers swap every step, so an even count leaves
// the roles as they were.)  Graphs are captured on a stream of the library's
// own, since the caller's may be the null stream, which cannot capture; they
// replay on the caller's stream.
struct RunGraph {
    int dev;
    pfs_field f[4];
    float dt, visc;
    int iters, tile, steps;
    float *ws;
    hipGraphExec_t exec;
    bool same(const RunGraph &o) const {
        if (o.dev != dev || o.dt != dt || o.visc != visc || o.iters != iters || o.tile != tile ||
            o.steps != steps || o.ws != ws)
            return false;
        for (int i = 0; i < 4; ++i)
            if (f[i].x != o.f[i].x || f[i].y != o.f[i].y || f[i].z != o.f[i].z ||
                f[i].data != o.f[i].data)
                return false;
        return true;
    }
};
static std::mutex g_graph_mutex;
static std::deque<RunGraph> g_graphs;         // most recently used last
static std::map<int, hipStream_t> g_capture;  // per device
constexpr int kMaxRunGraphs = 16;             // per device
constexpr int kRunBlock = 16;                 // steps per graph of a long run

// Records the launches of r.steps fused timesteps on `cs` (capturing).
static int capture_steps(const RunGraph &r, bool dye_fused, hipStream_t cs) {
    const pfs_field &vp = r.f[0], &tmp = r.f[1];
    pfs_field im = r.f[2], itm = r.f[3];
    for (int k = 0; k < r.steps; ++k) {
        StepArgs a{vp.data, tmp.data, dye_fused ? im.data : nullptr,
                   dye_fused ? itm.data : nullptr, vp.x, vp.y, r.dt, r.visc, r.iters, r.tile};
        PFS_TRY(fused_step(a, r.ws, cs, nullptr, 1));
        if (!dye_fused) PFS_TRY(pfs_advect_color(&im, &itm, &vp, r.dt, cs));
        std::swap(im.data, itm.data);
    }
    return PFS_OK;
}

// The cached graph equal to `want`, captured now if there is none.  Called
// with g_graph_mutex held.
static int run_graph(RunGraph want, bool dye_fused, hipGraphExec_t *out) {
    for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
        if (it->same(want)) {
            const RunGraph r = *it;
            g_graphs.erase(it);
            g_graphs.push_back(r);
            *out = r.exec;
            return PFS_OK;
        }
    hipStream_t &cs = g_capture[want.dev];
    if (!cs) PFS_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    PFS_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = capture_steps(want, dye_fused, cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    if (rc == PFS_OK && e != hipSuccess)
        rc = set_error(PFS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (rc == PFS_OK && hipGraphInstantiate(&want.exec, g, nullptr, nullptr, 0) != hipSuccess)
        rc = set_error(PFS_ERR_HIP, "hipGraphInstantiate failed");
    if (g) (void)hipGraphDestroy(g);
    if (rc != PFS_OK) return rc;
    int held = 0;
    for (const RunGraph &r : g_graphs) held += r.dev == want.dev;
    if (held >= kMaxRunGraphs) {
        // Evict this device's least recently used graph once nothing queued
        // on the device can still be running it.
        PFS_HIP_TRY(hipDeviceSynchronize());
        for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
            if (it->dev == want.dev) {
                (void)hipGraphExecDestroy(it->exec);
                g_graphs.erase(it);
                break;
            }
    }
    g_graphs.push_back(want);
    *out = want.exec;
    return PFS_OK;
}

int pfs_fluid_run(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                  float viscosity, int iters, int nsteps, void *stream) {
    if (!image || !itmp) return set_error(PFS_ERR_INVALID, "image or itmp is NULL");
    if (nsteps < 0) return set_error(PFS_ERR_INVALID, "nsteps < 0");
    PFS_TRY(check_step_fields(vp, tmp, iters));
    if (nsteps < 4 || !fused_ok(vp, tmp, iters)) {
        for (int i = 0; i < nsteps; ++i)
            PFS_TRY(timestep(vp, tmp, image, itmp, dt, viscosity, iters, stream, nullptr));
        return PFS_OK;
    }
    // One eager step checks the arguments and sizes the stream's workspace
    // outside any capture; an odd count takes a second one.
    PFS_TRY(timestep(vp, tmp, image, itmp, dt, viscosity, iters, stream, nullptr));
    int left = nsteps - 1;
    if (left % 2) {
        PFS_TRY(timestep(vp, tmp, image, itmp, dt, viscosity, iters, stream, nullptr));
        --left;
    }
    hipStream_t s = (hipStream_t)stream;
    RunGraph want{0, {*vp, *tmp, *image, *itmp}, dt, viscosity, iters, 0, 0, nullptr, nullptr};
    PFS_HIP_TRY(hipGetDevice(&want.dev));
    PFS_TRY(get_workspace(fused_workspace_bytes(vp->x, vp->y), s, &want.ws));  // allocated already
    const bool dye_fused = image->x == vp->x && image->y == vp->y && image->z == 4 &&
                           aligned16(image->data) && aligned16(itmp->data);
    want.tile = step_tile(vp->x, vp->y, iters, dye_fused);

    std::lock_guard<std::mutex> lock(g_graph_mutex);
    hipGraphExec_t exec = nullptr;
    if (left >= kRunBlock) {  // blocks of kRunBlock steps, then pairs
        want.steps = kRunBlock;
        PFS_TRY(run_graph(want, dye_fused, &exec));
        for (; left >= kRunBlock; left -= kRunBlock) PFS_HIP_TRY(hipGraphLaunch(exec, s));
    }
    if (left) {
        want.steps = 2;
        PFS_TRY(run_graph(want, dye_fused, &exec));
        for (; left; left -= 2) PFS_HIP_TRY(hipGraphLaunch(exec, s));
    }
    return PFS_OK;
}

int pfs_simulate_fluid_step(pfs_field *vp, pfs_field *tmp, float dt, float viscosity, int iters,
                            void *stream) {
    return timestep(vp, tmp, nullptr, nullptr, dt, viscosity, iters, stream, nullptr);
}

int pfs_fluid_timestep(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                       float viscosity, int iters, void *stream) {
    if (!image || !itmp) return set_error(PFS_ERR_INVALID, "image or itmp is NULL");
    return timestep(vp, tmp, image, itmp, dt, viscosity, iters, stream, nullptr);
}

// ---- the timestep that allocates nothing --------------------------------------
// The fused step's workspace planes (pressure start, divergence, E and, past
// 8 sweeps, the iterate planes: scene_ws_floats) in a file-scope device
// array instead of a hipMalloc'd block.  Each GPU the library runs on holds
// its own copy, at its own address, until the process ends.  Host code uses
// the address hipGetSymbolAddress gives for the current device, never the
// array's name (a host-side stand-in).  Nothing in it needs clearing: every
// plane is written by one kernel before the next reads it.
static __device__ __align__(256) float scratch[PFS_NOALLOC_SCRATCH_BYTES / sizeof(float)];

static std::mutex g_scratch_mutex;
static std::map<int, float *> g_scratch;  // device -> address of its `scratch`

static int scratch_here(float **out) {
    int dev = 0;
    PFS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch.find(dev);
    if (it == g_scratch.end()) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return set_error(PFS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(e));
        }
        it = g_scratch.emplace(dev, static_cast<float *>(p)).first;
    }
    *out = it->second;
    return PFS_OK;
}

int pfs_fluid_timestep_noalloc(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp,
                               float dt, float viscosity, int iters, void *stream) {
    if (!image || !itmp) return set_error(PFS_ERR_INVALID, "image or itmp is NULL");
    float *ws = nullptr;
    PFS_TRY(scratch_here(&ws));
    return timestep(vp, tmp, image, itmp, dt, viscosity, iters, stream, nullptr, 1, 0, ws,
                    PFS_NOALLOC_SCRATCH_BYTES);
}

int pfs_advect_color_step(pfs_field *image, pfs_field *itmp, const pfs_field *vp, float dt,
                          void *stream) {
    PFS_TRY(pfs_advect_color(image, itmp, vp, dt, stream));
    std::swap(image->data, itmp->data);  // src/fluid.cpp:317-319
    return PFS_OK;
}

int pfs_simulate_fluid_step_dev(float **vp, float **tmp, float dt, float viscosity, int vx,
                                int vy, int vz, int iters, void *stream) {
    if (!vp || !tmp) return set_error(PFS_ERR_INVALID, "vp or tmp is NULL");
    pfs_field a{vx, vy, vz, *vp}, b{vx, vy, vz, *tmp};
    PFS_TRY(pfs_simulate_fluid_step(&a, &b, dt, viscosity, iters, stream));
    *vp = a.data;  // identities are restored by the step
    *tmp = b.data;
    return PFS_OK;
}

int pfs_advect_color_step_dev(float **image, float **itmp, float **vp, float dt, int ix, int iy,
                              int iz, int vx, int vy, int vz, void *stream) {
    if (!image || !itmp || !vp) return set_error(PFS_ERR_INVALID, "NULL pointer argument");
    pfs_field im{ix, iy, iz, *image}, it{ix, iy, iz, *itmp}, v{vx, vy, vz, *vp};
    PFS_TRY(pfs_advect_color_step(&im, &it, &v, dt, stream));
    *image = im.data;
    *itmp = it.data;
    return PFS_OK;
}

int pfs_set_step_path(int path) {
    if (path != PFS_PATH_AUTO && path != PFS_PATH_PER_OP)
        return set_error(PFS_ERR_INVALID, "unknown step path");
    g_step_path = path;
    return PFS_OK;
}

int pfs_get_step_path(void) { return g_step_path; }

int pfs_set_step_tile(int side) {
    if (side != 0 && !valid_tile_side(side))
        return set_error(PFS_ERR_INVALID, "tile side must be 0 (automatic), 16, 32 or 64");
    g_step_tile = side;
    return PFS_OK;
}

int pfs_get_step_tile(int vx, int vy, int iters, int dye_fused) {
    if (vx < 1 || vy < 1 || iters < 1) return set_error(PFS_ERR_INVALID, "bad step shape");
    return step_tile(vx, vy, iters, dye_fused != 0);
}

int pfs_autotune_step(const pfs_field *vp, const pfs_field *tmp, const pfs_field *image,
                      const pfs_field *itmp, float dt, float viscosity, int iters, int reps,
                      void *stream, int *side, float *ms, int n_ms) {
    PFS_TRY(check_step_fields(vp, tmp, iters));
    if (reps < 1 || !side) return set_error(PFS_ERR_INVALID, "autotune: reps >= 1 and side needed");
    if (ms && n_ms < 3) return set_error(PFS_ERR_INVALID, "autotune: ms needs room for 3 values");
    if ((image == nullptr) != (itmp == nullptr))
        return set_error(PFS_ERR_INVALID, "image and itmp must both be given or both be NULL");
    if (image) {
        PFS_TRY(check_field(image, "image", 1));
        if (!same_shape(image, itmp)) return set_error(PFS_ERR_INVALID, "itmp shape != image shape");
    }
    if (!fused_ok(vp, tmp, iters))
        return set_error(PFS_ERR_UNSUPPORTED, "autotune: the step does not take the fused path");
    hipStream_t s = (hipStream_t)stream;
    // Scratch copies: the caller's buffers are read, never advanced.
    const size_t fb = (size_t)vp->x * vp->y * vp->z * sizeof(float);
    const size_t ib = image ? (size_t)image->x * image->y * image->z * sizeof(float) : 0;
    char *scratch = nullptr;
    const size_t fb_al = (fb + 255) & ~(size_t)255, ib_al = (ib + 255) & ~(size_t)255;
    if (hipMalloc((void **)&scratch, 2 * fb_al + 2 * ib_al) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(PFS_ERR_NOMEM, "autotune: scratch allocation failed");
    }
    pfs_field v{vp->x, vp->y, vp->z, reinterpret_cast<float *>(scratch)};
    pfs_field t{vp->x, vp->y, vp->z, reinterpret_cast<float *>(scratch + fb_al)};
    pfs_field im{}, it{};
    if (image) {
        im = pfs_field{image->x, image->y, image->z, reinterpret_cast<float *>(scratch + 2 * fb_al)};
        it = pfs_field{image->x, image->y, image->z,
                       reinterpret_cast<float *>(scratch + 2 * fb_al + ib_al)};
    }
    hipEvent_t ev[2] = {};
    int rc = PFS_OK;
    float best_ms = 0.0f;
    int best = 0;
    // The shape key as the caller's own step would see it (timestep's test).
    const bool dye_fused = image && image->x == vp->x && image->y == vp->y && image->z == 4 &&
                           aligned16(image->data) && aligned16(itmp->data);
    for (int k = 0; k < 2 && rc == PFS_OK; ++k)
        if (hipEventCreate(&ev[k]) != hipSuccess) rc = set_error(PFS_ERR_HIP, "autotune: event");
    for (int k = 0; k < 3 && rc == PFS_OK; ++k) {
        const int ts = kTileSides[k];
        hipError_t e = hipMemcpyAsync(v.data, vp->data, fb, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(t.data, tmp->data, fb, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && image)
            e = hipMemcpyAsync(im.data, image->data, ib, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            rc = set_error(PFS_ERR_HIP, std::string("autotune copy: ") + hipGetErrorString(e));
            break;
        }
        // One untimed step (first launches of these kernels), then reps timed.
        for (int r = 0; r <= reps && rc == PFS_OK; ++r) {
            if (r == 1 && hipEventRecord(ev[0], s) != hipSuccess) rc = set_error(PFS_ERR_HIP, "autotune: event");
            if (rc == PFS_OK)
                rc = timestep(&v, &t, image ? &im : nullptr, image ? &it : nullptr, dt, viscosity,
                              iters, stream, nullptr, 1, ts);
        }
        if (rc != PFS_OK) break;
        float el = 0.0f;
        e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&el, ev[0], ev[1]);
        if (e != hipSuccess) {
            rc = set_error(PFS_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
            break;
        }
        el /= (float)reps;
        if (ms) ms[k] = el;
        if (best == 0 || el < best_ms) {
            best = ts;
            best_ms = el;
        }
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(s);
    (void)hipFree(scratch);
    if (rc != PFS_OK) return rc;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tuned[{vp->x, vp->y, iters, dye_fused ? 1 : 0}] = best;
    }
    *side = best;
    return PFS_OK;
}

int pfs_clear_step_tuning(void) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned.clear();
    return PFS_OK;
}

size_t pfs_step_workspace_bytes(int vx, int vy) {
    if (vx < 1 || vy < 1) return 0;
    return fused_workspace_bytes(vx, vy);
}

int pfs_profile_step(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                     float viscosity, int iters, int reps, void *stream, float *ms, int n_ms) {
    if (!ms || n_ms < 2) return set_error(PFS_ERR_INVALID, "ms needs room for 2 values");
    if (reps < 1) return set_error(PFS_ERR_INVALID, "reps must be >= 1");
    if ((image == nullptr) != (itmp == nullptr))
        return set_error(PFS_ERR_INVALID, "image and itmp must both be given or both be NULL");
    hipEvent_t ev[4] = {};
    for (auto &e : ev) PFS_HIP_TRY(hipEventCreate(&e));
    int rc = timestep(vp, tmp, image, itmp, dt, viscosity, iters, stream, ev, reps);
    if (rc == PFS_OK) {
        hipError_t e = hipEventSynchronize(ev[3]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms[0], ev[0], ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms[1], ev[2], ev[3]);
        ms[0] /= (float)reps;
        ms[1] /= (float)reps;
        if (e != hipSuccess)
            rc = set_error(PFS_ERR_HIP, std::string("event timing: ") + hipGetErrorString(e));
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int pfs_debug_advect_keys(const pfs_field *vp, float dt, int32_t *keys, void *stream) {
    PFS_TRY(check_field(vp, "vp", 2));
    if (!keys) return set_error(PFS_ERR_INVALID, "keys is NULL");
    return launch_advect_keys(vp->data, keys, vp->x, vp->y, vp->z, dt, (hipStream_t)stream);
}

int pfs_debug_advect_color_keys(const pfs_field *image, const pfs_field *vp, float dt,
                                int32_t *keys, void *stream) {
    PFS_TRY(check_field(image, "image", 1));
    PFS_TRY(check_field(vp, "vp", 2));
    if (!keys) return set_error(PFS_ERR_INVALID, "keys is NULL");
    return launch_advect_color_keys(vp->data, keys, image->x, image->y, vp->x, vp->y, vp->z, dt,
                                    (hipStream_t)stream);
}

}  // extern "C"
