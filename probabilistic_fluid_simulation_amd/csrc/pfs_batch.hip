// Batched fluid step: many independent scenes advanced together.
//
// The reference advances one scene per call (src/main.cpp:219-240); small
// scenes leave most of the GPU idle and are bound by launch latency.  A
// pfs_batch holds a table of scenes (caller-owned buffers, batch-owned
// workspaces) and advances every scene by one timestep — the reference's
// simulate_fluid_step + advect_color_step — with one launch per kernel
// stage for all scenes together:
//
//   k_diffuse_batch   every tile of every scene (scenes grouped by K)
//   k_pressure_batch  every tile of every scene, dye fused where the dye has
//                     the field's size
//   k_color_batch     the dye of the remaining scenes
//
// Each workgroup finds its scene through a tile -> scene table and runs the
// same per-tile body as the single-scene kernels (pfs_step_kernels.h), so
// every scene's result is bit-identical to stepping it alone.  The launch
// sequences of 8 and of 2 consecutive steps (the dye buffers swap every step)
// are captured once in hipGraphs and replayed while membership is unchanged;
// adding or removing a scene rebuilds the tables and the graphs.  In a
// captured run the dye pass of step k is a branch of its own: it reads the
// velocity that step k's pressure kernels wrote and writes only the dye, so
// it runs beside step k+1's diffusion kernels (which only read the
// velocity); step k+1's pressure kernels, which overwrite it, wait for it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pfs.h"
#include "pfs_internal.h"
#include "pfs_step_kernels.h"

namespace pfs {
namespace {

constexpr int kCBX = 64, kCBY = 4;  // dye threads per block: 64 x 4
constexpr int kCRows = 4;           // pixels per thread, kCBY rows apart: 64 x 16 per block

template <int H, bool LAST, int TS>
__global__ __launch_bounds__((Geo<H, TS>::NT)) void k_diffuse_batch(
    const DiffuseArgs *__restrict__ tab, const int *__restrict__ tile_scene, int ntiles) {
    const int b = xcd_tile(blockIdx.x, ntiles);
    const DiffuseArgs a = tab[tile_scene[b]];
    diffuse_tile<H, LAST, TS>(a, b - a.tile_base);
}

template <int H, bool LAST, int TS>
__global__ __launch_bounds__((Geo<H, TS>::NT)) void k_pressure_batch(
    const PressureArgs *__restrict__ tab, const int *__restrict__ tile_scene, int ntiles) {
    const int b = xcd_tile(blockIdx.x, ntiles);
    const PressureArgs a = tab[tile_scene[b]];
    pressure_tile<H, LAST, TS>(a, b - a.tile_base);
}

__global__ __launch_bounds__(kCBX *kCBY) void k_color_batch(const ColorArgs *__restrict__ tab,
                                                            const int *__restrict__ block_scene) {
    const ColorArgs a = tab[block_scene[blockIdx.x]];
    const int blk = blockIdx.x - a.block_base;
    const int i = (blk % a.blocks_x) * kCBX + threadIdx.x % kCBX;
    const int j = (blk / a.blocks_x) * (kCBY * kCRows) + threadIdx.x / kCBX;
    advect_color_column<kCRows>(a, i, j, kCBY);
}

using DiffuseBatchFn = void (*)(const DiffuseArgs *, const int *, int);
using PressureBatchFn = void (*)(const PressureArgs *, const int *, int);

#define PFS_H_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9)

template <int TS>
DiffuseBatchFn diffuse_batch_for_ts(int H, bool last) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return last ? k_diffuse_batch<h, true, TS> : k_diffuse_batch<h, false, TS>;
        PFS_H_CASES(PFS_CASE)
#undef PFS_CASE
    }
    return nullptr;
}

template <int TS>
PressureBatchFn pressure_batch_for_ts(int H, bool last) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return last ? k_pressure_batch<h, true, TS> : k_pressure_batch<h, false, TS>;
        PFS_H_CASES(PFS_CASE)
#undef PFS_CASE
    }
    return nullptr;
}

DiffuseBatchFn diffuse_batch_for(int H, bool last, int ts) {
    return ts == 16 ? diffuse_batch_for_ts<16>(H, last)
                    : ts == 32 ? diffuse_batch_for_ts<32>(H, last) : diffuse_batch_for_ts<64>(H, last);
}

PressureBatchFn pressure_batch_for(int H, bool last, int ts) {
    return ts == 16 ? pressure_batch_for_ts<16>(H, last)
                    : ts == 32 ? pressure_batch_for_ts<32>(H, last)
                               : pressure_batch_for_ts<64>(H, last);
}

constexpr int kMaxPlan = 64;  // chunks of one step: K up to 512

}  // namespace
}  // namespace pfs

using namespace pfs;

struct pfs_batch {
    struct Scene {
        bool live = false;
        pfs_field vp{}, tmp{};
        pfs_field dye[2]{};  // dye[(cur0 + steps) & 1] holds the current image
        int cur0 = 0;
        long long t0 = 0;  // batch step count when the scene joined
        float dt = 0, visc = 0;
        int iters = 0;
        float *ws = nullptr;
        size_t ws_bytes = 0;
        bool dye_fused = false;
        long long steps(long long t) const { return t - t0; }
        int cur(long long t) const { return (int)((cur0 + t - t0) & 1); }
    };
    // Tables of one K group, all in device memory.
    struct Group {
        int iters = 0, nchunks = 0, ntiles = 0;
        Chunk plan[kMaxPlan];
        int *tile_scene = nullptr;
        DiffuseArgs *dtab[kMaxPlan] = {};
        PressureArgs *ptab[2][kMaxPlan] = {};
    };
    // Device memory or a graph that queued work may still use: released once
    // an event recorded after that work has completed.
    struct Retired {
        hipEvent_t ev = nullptr;
        void *mem = nullptr;
        hipGraphExec_t graph = nullptr;
    };

    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<Scene> scenes;
    bool dirty = true;
    bool use_graph = true;
    int tile_forced = 0;  // 0: the rule over the live scenes' tile counts
    int tile = kTile;     // tile side of the current tables
    std::vector<Group> groups;
    ColorArgs *ctab[2] = {};
    int *block_scene = nullptr;
    int nblocks = 0;
    void *dev_tables = nullptr;
    static constexpr int kRunSteps = 8;  // steps of the longer captured run (even)
    hipGraphExec_t graph = nullptr;      // 2 steps
    hipGraphExec_t graph_run = nullptr;  // kRunSteps steps
    hipStream_t side = nullptr;          // the dye branch while capturing
    hipEvent_t fork_ev[kRunSteps] = {}, join_ev[kRunSteps] = {};
    long long t = 0;        // steps enqueued so far
    long long t_build = 0;  // t when the tables were built
    std::vector<Retired> retired;

    // Frees retired items whose work has finished (all of them if `wait`).
    int reap(bool wait) {
        std::vector<Retired> keep;
        for (auto &r : retired) {
            hipError_t e = wait ? hipEventSynchronize(r.ev) : hipEventQuery(r.ev);
            if (e == hipErrorNotReady) {
                keep.push_back(r);
                continue;
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                return set_error(PFS_ERR_HIP, std::string("batch event: ") + hipGetErrorString(e));
            }
            if (r.graph) (void)hipGraphExecDestroy(r.graph);
            if (r.mem) (void)hipFree(r.mem);
            (void)hipEventDestroy(r.ev);
        }
        retired.swap(keep);
        return PFS_OK;
    }

    int retire(void *mem, hipGraphExec_t g) {
        if (!mem && !g) return PFS_OK;
        Retired r;
        PFS_HIP_TRY(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
        PFS_HIP_TRY(hipEventRecord(r.ev, stream));
        r.mem = mem;
        r.graph = g;
        retired.push_back(r);
        return PFS_OK;
    }

    int sync() {
        PFS_HIP_TRY(hipStreamSynchronize(stream));
        return reap(true);
    }

    int drop_tables() {
        PFS_TRY(retire(dev_tables, graph));
        PFS_TRY(retire(nullptr, graph_run));
        dev_tables = nullptr;
        graph = nullptr;
        graph_run = nullptr;
        groups.clear();
        ctab[0] = ctab[1] = nullptr;
        block_scene = nullptr;
        nblocks = 0;
        dirty = true;
        return PFS_OK;
    }

    // The rule over every live scene's tiles (one launch covers them all).
    int auto_tile() const {
        long long n64 = 0, n32 = 0;
        for (const Scene &s : scenes)
            if (s.live) {
                n64 += tile_count(s.vp.x, s.vp.y, 64);
                n32 += tile_count(s.vp.x, s.vp.y, 32);
            }
        return tile_for_counts(n64, n32);
    }

    // Rebuilds every table for the live scenes.  Table parity q is the
    // parity of (t - t_build): scene s then reads its image from
    // dye[(cur0 + t_build - t0 + q) & 1].
    int rebuild() {
        PFS_TRY(reap(false));
        PFS_TRY(drop_tables());
        t_build = t;
        tile = tile_forced ? tile_forced : auto_tile();
        std::map<int, std::vector<int>> by_k;
        std::vector<int> unfused;
        for (int i = 0; i < (int)scenes.size(); ++i) {
            if (!scenes[i].live) continue;
            by_k[scenes[i].iters].push_back(i);
            if (!scenes[i].dye_fused) unfused.push_back(i);
        }
        std::vector<char> img;
        auto reserve = [&](size_t bytes) {
            size_t off = (img.size() + 255) & ~(size_t)255;
            img.resize(off + bytes);
            return off;
        };
        struct Off {
            size_t tile_scene, dtab[kMaxPlan], ptab[2][kMaxPlan];
        };
        std::vector<Off> offs;
        for (auto &kv : by_k) {
            Group g;
            g.iters = kv.first;
            g.nchunks = chunk_plan(g.iters, g.plan, kMaxPlan);
            if (!g.plan[g.nchunks - 1].last)
                return set_error(PFS_ERR_UNSUPPORTED, "iters too large for the batch");
            const std::vector<int> &ids = kv.second;
            std::vector<int> tile_scene;
            std::vector<DiffuseArgs> d(ids.size());
            std::vector<PressureArgs> p(ids.size());
            for (size_t k = 0; k < ids.size(); ++k) {
                const Scene &s = scenes[ids[k]];
                d[k] = diffuse_args(s.vp.data, s.tmp.data, s.vp.x, s.vp.y, s.dt, s.visc, s.iters,
                                    s.ws, tile);
                d[k].tile_base = (int)tile_scene.size();
                p[k] = pressure_args(s.vp.data, s.tmp.data, nullptr, nullptr, s.vp.x, s.vp.y, s.dt,
                                     s.ws, tile);
                p[k].tile_base = d[k].tile_base;
                tile_scene.insert(tile_scene.end(), d[k].ntiles, (int)k);
            }
            g.ntiles = (int)tile_scene.size();
            Off o{};
            o.tile_scene = reserve(tile_scene.size() * sizeof(int));
            std::memcpy(img.data() + o.tile_scene, tile_scene.data(), tile_scene.size() * sizeof(int));
            for (int c = 0; c < g.nchunks; ++c) {
                o.dtab[c] = reserve(d.size() * sizeof(DiffuseArgs));
                for (size_t k = 0; k < ids.size(); ++k) {
                    DiffuseArgs a = d[k];
                    diffuse_chunk(a, c, g.plan[c], scenes[ids[k]].ws);
                    std::memcpy(img.data() + o.dtab[c] + k * sizeof(DiffuseArgs), &a, sizeof(a));
                }
                for (int q = 0; q < 2; ++q) {
                    o.ptab[q][c] = reserve(d.size() * sizeof(PressureArgs));
                    for (size_t k = 0; k < ids.size(); ++k) {
                        const Scene &s = scenes[ids[k]];
                        const int cur = (int)((s.cur(t_build) + q) & 1);
                        PressureArgs a = p[k];
                        pressure_chunk(a, c, g.plan[c], s.ws, s.dye_fused ? s.dye[cur].data : nullptr,
                                       s.dye_fused ? s.dye[cur ^ 1].data : nullptr);
                        std::memcpy(img.data() + o.ptab[q][c] + k * sizeof(PressureArgs), &a,
                                    sizeof(a));
                    }
                }
            }
            groups.push_back(g);
            offs.push_back(o);
        }
        size_t c_off[2] = {0, 0}, bs_off = 0;
        if (!unfused.empty()) {
            std::vector<int> block_scene_h;
            std::vector<ColorArgs> c[2];
            for (size_t k = 0; k < unfused.size(); ++k) {
                const Scene &s = scenes[unfused[k]];
                for (int q = 0; q < 2; ++q) {
                    const int cur = (int)((s.cur(t_build) + q) & 1);
                    const pfs_field &im = s.dye[cur];
                    ColorArgs a = color_args(im.data, s.dye[cur ^ 1].data, s.vp.data, im.x, im.y,
                                             im.z, s.vp.x, s.vp.y, s.vp.z, s.dt);
                    a.blocks_x = (im.x + kCBX - 1) / kCBX;
                    a.block_base = (int)block_scene_h.size();
                    c[q].push_back(a);
                }
                const int nb =
                    c[0].back().blocks_x * ((s.dye[0].y + kCBY * kCRows - 1) / (kCBY * kCRows));
                block_scene_h.insert(block_scene_h.end(), nb, (int)k);
            }
            nblocks = (int)block_scene_h.size();
            bs_off = reserve(block_scene_h.size() * sizeof(int));
            std::memcpy(img.data() + bs_off, block_scene_h.data(), block_scene_h.size() * sizeof(int));
            for (int q = 0; q < 2; ++q) {
                c_off[q] = reserve(c[q].size() * sizeof(ColorArgs));
                std::memcpy(img.data() + c_off[q], c[q].data(), c[q].size() * sizeof(ColorArgs));
            }
        }
        if (!img.empty()) {
            if (hipMalloc(&dev_tables, img.size()) != hipSuccess) {
                (void)hipGetLastError();
                dev_tables = nullptr;
                return set_error(PFS_ERR_NOMEM, "batch tables: device allocation failed");
            }
            // Synchronous: the new block is not yet used by any queued work, and
            // the copy has completed (and read `img`) when the call returns.
            PFS_HIP_TRY(hipMemcpy(dev_tables, img.data(), img.size(), hipMemcpyHostToDevice));
        }
        char *base = static_cast<char *>(dev_tables);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            Group &g = groups[gi];
            g.tile_scene = reinterpret_cast<int *>(base + offs[gi].tile_scene);
            for (int c = 0; c < g.nchunks; ++c) {
                g.dtab[c] = reinterpret_cast<DiffuseArgs *>(base + offs[gi].dtab[c]);
                for (int q = 0; q < 2; ++q)
                    g.ptab[q][c] = reinterpret_cast<PressureArgs *>(base + offs[gi].ptab[q][c]);
            }
        }
        if (nblocks) {
            block_scene = reinterpret_cast<int *>(base + bs_off);
            for (int q = 0; q < 2; ++q) ctab[q] = reinterpret_cast<ColorArgs *>(base + c_off[q]);
        }
        dirty = false;
        return PFS_OK;
    }

    // One timestep of every scene, with the tables of parity q.
    int enqueue_step(int q) {
        for (const Group &g : groups) {
            for (int c = 0; c < g.nchunks; ++c) {
                hipLaunchKernelGGL(diffuse_batch_for(g.plan[c].H, g.plan[c].last, tile),
                                   dim3(g.ntiles), dim3(threads_for(g.plan[c].H, tile)), 0, stream,
                                   g.dtab[c], g.tile_scene, g.ntiles);
                PFS_TRY(check_launch("batched diffuse"));
            }
            for (int c = 0; c < g.nchunks; ++c) {
                hipLaunchKernelGGL(pressure_batch_for(g.plan[c].H, g.plan[c].last, tile),
                                   dim3(g.ntiles), dim3(threads_for(g.plan[c].H, tile)), 0, stream,
                                   g.ptab[q][c], g.tile_scene, g.ntiles);
                PFS_TRY(check_launch("batched pressure"));
            }
        }
        if (nblocks) {
            hipLaunchKernelGGL(k_color_batch, dim3(nblocks), dim3(kCBX * kCBY), 0, stream,
                               ctab[q], block_scene);
            PFS_TRY(check_launch("batched dye"));
        }
        return PFS_OK;
    }

    // The side stream and events of a captured run (made once; on failure
    // nothing is kept and the next capture tries again).
    int capture_resources() {
        if (side) return PFS_OK;
        hipError_t e = hipSuccess;
        for (int k = 0; k < kRunSteps && e == hipSuccess; ++k) {
            e = hipEventCreateWithFlags(&fork_ev[k], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&join_ev[k], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            side = nullptr;
            free_capture_resources();
            return set_error(PFS_ERR_HIP, std::string("batch capture resources: ") + hipGetErrorString(e));
        }
        return PFS_OK;
    }

    void free_capture_resources() {
        for (int k = 0; k < kRunSteps; ++k) {
            if (fork_ev[k]) (void)hipEventDestroy(fork_ev[k]);
            if (join_ev[k]) (void)hipEventDestroy(join_ev[k]);
            fork_ev[k] = join_ev[k] = nullptr;
        }
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
    }

    // n steps (n even, table parity k & 1 for step k), the dye pass of each
    // on the side stream: forked after the step's pressure kernels, joined
    // before the next step's pressure kernels and at the end.
    int enqueue_run(int n) {
        for (int k = 0; k < n; ++k) {
            const int q = k & 1;
            bool joined = k == 0 || nblocks == 0;
            for (const Group &g : groups) {
                for (int c = 0; c < g.nchunks; ++c) {
                    hipLaunchKernelGGL(diffuse_batch_for(g.plan[c].H, g.plan[c].last, tile),
                                       dim3(g.ntiles), dim3(threads_for(g.plan[c].H, tile)), 0,
                                       stream, g.dtab[c], g.tile_scene, g.ntiles);
                    PFS_TRY(check_launch("batched diffuse"));
                }
                if (!joined) {
                    PFS_HIP_TRY(hipStreamWaitEvent(stream, join_ev[k - 1], 0));
                    joined = true;
                }
                for (int c = 0; c < g.nchunks; ++c) {
                    hipLaunchKernelGGL(pressure_batch_for(g.plan[c].H, g.plan[c].last, tile),
                                       dim3(g.ntiles), dim3(threads_for(g.plan[c].H, tile)), 0,
                                       stream, g.ptab[q][c], g.tile_scene, g.ntiles);
                    PFS_TRY(check_launch("batched pressure"));
                }
            }
            if (nblocks) {
                PFS_HIP_TRY(hipEventRecord(fork_ev[k], stream));
                PFS_HIP_TRY(hipStreamWaitEvent(side, fork_ev[k], 0));
                hipLaunchKernelGGL(k_color_batch, dim3(nblocks), dim3(kCBX * kCBY), 0, side, ctab[q],
                                   block_scene);
                PFS_TRY(check_launch("batched dye"));
                PFS_HIP_TRY(hipEventRecord(join_ev[k], side));
            }
        }
        if (nblocks) PFS_HIP_TRY(hipStreamWaitEvent(stream, join_ev[n - 1], 0));
        return PFS_OK;
    }

    int capture(int n, hipGraphExec_t *out) {
        PFS_TRY(capture_resources());
        hipGraph_t g = nullptr;
        PFS_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_run(n);
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc != PFS_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return set_error(PFS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            *out = nullptr;
            return set_error(PFS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        return PFS_OK;
    }

    // Captures the graphs the next steps would capture on first use.
    int prepare_graphs() {
        if (dirty) PFS_TRY(rebuild());
        if (!use_graph) return PFS_OK;
        if (!graph) PFS_TRY(capture(2, &graph));
        if (!graph_run) PFS_TRY(capture(kRunSteps, &graph_run));
        return PFS_OK;
    }

    int step(int n) {
        if (n < 0) return set_error(PFS_ERR_INVALID, "negative step count");
        if (dirty) PFS_TRY(rebuild());
        bool any = false;
        for (auto &s : scenes) any = any || s.live;
        if (!any) return PFS_OK;
        for (int left = n; left > 0;) {
            const int q = (int)((t - t_build) & 1);
            if (use_graph && q == 0 && left >= kRunSteps) {
                if (!graph_run) PFS_TRY(capture(kRunSteps, &graph_run));
                PFS_HIP_TRY(hipGraphLaunch(graph_run, stream));
                t += kRunSteps;
                left -= kRunSteps;
            } else if (use_graph && q == 0 && left >= 2) {
                if (!graph) PFS_TRY(capture(2, &graph));
                PFS_HIP_TRY(hipGraphLaunch(graph, stream));
                t += 2;
                left -= 2;
            } else {
                PFS_TRY(enqueue_step(q));
                t += 1;
                left -= 1;
            }
        }
        return PFS_OK;
    }
};

extern "C" {

int pfs_batch_create(pfs_batch **out, void *stream) {
    if (!out) return set_error(PFS_ERR_INVALID, "out is NULL");
    pfs_batch *b = new pfs_batch();
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return set_error(PFS_ERR_HIP, "hipStreamCreateWithFlags failed");
        }
        b->own_stream = true;
    }
    *out = b;
    return PFS_OK;
}

int pfs_batch_destroy(pfs_batch *b) {
    if (!b) return PFS_OK;
    int rc = b->drop_tables();
    if (rc == PFS_OK) rc = b->sync();
    for (auto &s : b->scenes)
        if (s.ws) (void)hipFree(s.ws);
    b->free_capture_resources();
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return rc;
}

int pfs_batch_add(pfs_batch *b, const pfs_field *vp, const pfs_field *tmp, const pfs_field *image,
                  const pfs_field *itmp, float dt, float viscosity, int iters, int *slot) {
    if (!b || !slot) return set_error(PFS_ERR_INVALID, "batch or slot is NULL");
    PFS_TRY(check_field(vp, "vp", 4));
    PFS_TRY(check_field(tmp, "tmp", 4));
    PFS_TRY(check_field(image, "image", 1));
    PFS_TRY(check_field(itmp, "itmp", 1));
    if (vp->x != tmp->x || vp->y != tmp->y || vp->z != tmp->z)
        return set_error(PFS_ERR_INVALID, "tmp shape != vp shape");
    if (image->x != itmp->x || image->y != itmp->y || image->z != itmp->z)
        return set_error(PFS_ERR_INVALID, "itmp shape != image shape");
    const float *ptrs[4] = {vp->data, tmp->data, image->data, itmp->data};
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j)
            if (ptrs[i] == ptrs[j]) return set_error(PFS_ERR_INVALID, "scene buffers alias");
    auto al = [](const void *p) { return (uintptr_t)p % 16 == 0; };
    if (vp->z != 4 || iters < 1 || iters > kMaxChunk * kMaxPlan || !al(vp->data) || !al(tmp->data))
        return set_error(PFS_ERR_UNSUPPORTED,
                         "batched scenes need 4-channel, 16-byte aligned fields and 1 <= iters <= 512");
    pfs_batch::Scene s;
    s.live = true;
    s.vp = *vp;
    s.tmp = *tmp;
    s.dye[0] = *image;
    s.dye[1] = *itmp;
    s.dt = dt;
    s.visc = viscosity;
    s.iters = iters;
    s.dye_fused = image->x == vp->x && image->y == vp->y && image->z == 4 && al(image->data) &&
                  al(itmp->data);
    s.t0 = b->t;
    s.ws_bytes = scene_ws_floats(vp->x, vp->y, iters) * sizeof(float);
    if (hipMalloc(&s.ws, s.ws_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(PFS_ERR_NOMEM, "scene workspace allocation failed");
    }
    // Reuse a free slot, else append.
    int id = -1;
    for (int i = 0; i < (int)b->scenes.size(); ++i)
        if (!b->scenes[i].live) {
            id = i;
            break;
        }
    if (id < 0) {
        id = (int)b->scenes.size();
        b->scenes.push_back(s);
    } else {
        b->scenes[id] = s;
    }
    b->dirty = true;
    *slot = id;
    return PFS_OK;
}

int pfs_batch_remove(pfs_batch *b, int slot) {
    if (!b || slot < 0 || slot >= (int)b->scenes.size() || !b->scenes[slot].live)
        return set_error(PFS_ERR_INVALID, "no such scene");
    // Queued steps may still use the workspace: it is released once they end.
    PFS_TRY(b->retire(b->scenes[slot].ws, nullptr));
    b->scenes[slot] = pfs_batch::Scene();
    b->dirty = true;
    return PFS_OK;
}

int pfs_batch_step(pfs_batch *b, int nsteps) {
    if (!b) return set_error(PFS_ERR_INVALID, "batch is NULL");
    return b->step(nsteps);
}

int pfs_batch_sync(pfs_batch *b) {
    if (!b) return set_error(PFS_ERR_INVALID, "batch is NULL");
    return b->sync();
}

int pfs_batch_scene(pfs_batch *b, int slot, pfs_field *image, pfs_field *itmp, long long *steps) {
    if (!b || slot < 0 || slot >= (int)b->scenes.size() || !b->scenes[slot].live)
        return set_error(PFS_ERR_INVALID, "no such scene");
    const pfs_batch::Scene &s = b->scenes[slot];
    const int cur = s.cur(b->t);
    if (image) *image = s.dye[cur];
    if (itmp) *itmp = s.dye[cur ^ 1];
    if (steps) *steps = s.steps(b->t);
    return PFS_OK;
}

int pfs_batch_count(pfs_batch *b) {
    if (!b) return 0;
    int n = 0;
    for (auto &s : b->scenes) n += s.live ? 1 : 0;
    return n;
}

size_t pfs_batch_workspace_bytes(pfs_batch *b) {
    if (!b) return 0;
    size_t n = 0;
    for (auto &s : b->scenes) n += s.live ? s.ws_bytes : 0;
    return n;
}

int pfs_batch_set_graphs(pfs_batch *b, int enable) {
    if (!b) return set_error(PFS_ERR_INVALID, "batch is NULL");
    b->use_graph = enable != 0;
    return PFS_OK;
}

void *pfs_batch_stream(pfs_batch *b) { return b ? (void *)b->stream : nullptr; }

int pfs_batch_set_tile(pfs_batch *b, int side) {
    if (!b) return set_error(PFS_ERR_INVALID, "batch is NULL");
    if (side != 0 && !valid_tile_side(side))
        return set_error(PFS_ERR_INVALID, "tile side must be 0 (automatic), 16, 32 or 64");
    if (side != b->tile_forced) {
        b->tile_forced = side;
        b->dirty = true;
    }
    return PFS_OK;
}

int pfs_batch_tile(pfs_batch *b) {
    if (!b) return set_error(PFS_ERR_INVALID, "batch is NULL");
    return b->tile_forced ? b->tile_forced : b->auto_tile();
}

int pfs_batch_autotune(pfs_batch *b, int reps, int *side, float *ms, int n_ms) {
    if (!b || !side || reps < 1) return set_error(PFS_ERR_INVALID, "autotune: batch, side, reps >= 1");
    if (ms && n_ms < 3) return set_error(PFS_ERR_INVALID, "autotune: ms needs room for 3 values");
    if (pfs_batch_count(b) == 0) return set_error(PFS_ERR_INVALID, "autotune: the batch is empty");
    PFS_TRY(b->sync());
    // Snapshot every live scene's four buffers; the timed steps run on the
    // scenes themselves and the snapshot is copied back after each side.
    struct Buf {
        float *dst;
        size_t bytes, off;
    };
    std::vector<Buf> bufs;
    size_t total = 0;
    for (auto &sc : b->scenes) {
        if (!sc.live) continue;
        for (const pfs_field *f : {&sc.vp, &sc.tmp, &sc.dye[0], &sc.dye[1]}) {
            const size_t bytes = (size_t)f->x * f->y * f->z * sizeof(float);
            bufs.push_back({f->data, bytes, total});
            total += (bytes + 255) & ~(size_t)255;
        }
    }
    char *snap = nullptr;
    if (hipMalloc((void **)&snap, total) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(PFS_ERR_NOMEM, "autotune: snapshot allocation failed");
    }
    auto copy_all = [&](bool save) -> int {
        for (const Buf &x : bufs) {
            hipError_t e = save ? hipMemcpyAsync(snap + x.off, x.dst, x.bytes, hipMemcpyDeviceToDevice, b->stream)
                                : hipMemcpyAsync(x.dst, snap + x.off, x.bytes, hipMemcpyDeviceToDevice, b->stream);
            if (e != hipSuccess)
                return set_error(PFS_ERR_HIP, std::string("autotune copy: ") + hipGetErrorString(e));
        }
        return PFS_OK;
    };
    const long long t_saved = b->t;
    const int forced_saved = b->tile_forced;
    hipEvent_t ev[2] = {};
    int rc = copy_all(true);
    for (int k = 0; k < 2 && rc == PFS_OK; ++k)
        if (hipEventCreate(&ev[k]) != hipSuccess) rc = set_error(PFS_ERR_HIP, "autotune: event");
    int best = 0;
    float best_ms = 0.0f;
    for (int k = 0; k < 3 && rc == PFS_OK; ++k) {
        b->tile_forced = kTileSides[k];
        b->dirty = true;
        rc = b->step(2);  // untimed: tables, graphs, first launches
        if (rc == PFS_OK) rc = b->prepare_graphs();
        if (rc == PFS_OK && hipEventRecord(ev[0], b->stream) != hipSuccess)
            rc = set_error(PFS_ERR_HIP, "autotune: event");
        if (rc == PFS_OK) rc = b->step(2 * reps);
        float el = 0.0f;
        if (rc == PFS_OK) {
            hipError_t e = hipEventRecord(ev[1], b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&el, ev[0], ev[1]);
            if (e != hipSuccess)
                rc = set_error(PFS_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
        }
        const int rc2 = copy_all(false);  // back to the state before tuning
        b->t = t_saved;
        b->dirty = true;
        if (rc == PFS_OK) rc = rc2;
        if (rc != PFS_OK) break;
        el /= (float)(2 * reps);
        if (ms) ms[k] = el;
        if (best == 0 || el < best_ms) {
            best = kTileSides[k];
            best_ms = el;
        }
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    const int rcs = b->sync();
    (void)hipFree(snap);
    b->tile_forced = rc == PFS_OK ? best : forced_saved;
    b->dirty = true;
    if (rc != PFS_OK) return rc;
    if (rcs != PFS_OK) return rcs;
    *side = best;
    return PFS_OK;
}

}  // extern "C"
