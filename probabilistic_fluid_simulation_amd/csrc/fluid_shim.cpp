// The reference's includes/fluid.hpp API (include/fluid.hpp here) on top of
// the C ABI.  vp_field arguments may hold host memory — computed on device
// mirrors of the caller's buffers (kept between calls), with what the
// operator reads copied up and what it writes copied back on the null
// stream, returning with the results in host memory like the reference's
// CPU build — or device memory, enqueued asynchronously like the
// reference's CUDA build.  Pointer swaps of the reference
// (src/fluid.cpp:188-194, 260-265, 317-319) are reproduced on the caller's
// structs in both cases.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "fluid.hpp"
#include "pfs.h"

namespace {

std::atomic<int> g_iters{NUM_JACOBI_ITERS};

void report(const char *fn, int rc) {
    if (rc != PFS_OK) std::fprintf(stderr, "pfs: %s failed (%d): %s\n", fn, rc, pfs_last_error());
}

bool is_device(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();  // plain host memory is "not registered"
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

size_t nbytes(const vp_field *f) { return sizeof(float) * (size_t)f->x * f->y * f->z; }

// Device mirrors of the caller's host buffers, kept between calls: the
// reference's CPU driver hands the same four buffers to every call (the
// structs swap pointers among them, src/main.cpp:219-240), so after the
// first timestep no call allocates.  A mirror is keyed by host address and
// size; at most kMirrors are kept, the least recently used one released
// first.  Every call synchronises the device before it returns, so no
// queued work uses a mirror between calls.
constexpr int kMirrors = 16;
struct Mirror {
    float *host;
    size_t bytes;
    float *dev;
    unsigned long long used;
};
std::mutex g_mirror_mutex;  // held for a whole host-flavour call
std::vector<Mirror> g_mirrors;
unsigned long long g_tick = 0;

float *mirror_of(float *host, size_t bytes) {
    for (auto &m : g_mirrors)
        if (m.host == host && m.bytes == bytes) {
            m.used = ++g_tick;
            return m.dev;
        }
    if ((int)g_mirrors.size() >= kMirrors) {
        auto lru = g_mirrors.begin();
        for (auto it = g_mirrors.begin(); it != g_mirrors.end(); ++it)
            if (it->used < lru->used) lru = it;
        (void)hipFree(lru->dev);
        g_mirrors.erase(lru);
    }
    // A mirror of the same host address but another size is replaced.
    for (auto it = g_mirrors.begin(); it != g_mirrors.end(); ++it)
        if (it->host == host) {
            (void)hipFree(it->dev);
            g_mirrors.erase(it);
            break;
        }
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_mirrors.push_back({host, bytes, static_cast<float *>(d), ++g_tick});
    return static_cast<float *>(d);
}

// Runs `op` on pfs_field views of the n caller fields.  Host fields go
// through their device mirrors: field i is uploaded first if bit i of `up`
// is set (the operator reads it, or leaves part of it untouched) and copied
// back afterwards if bit i of `down` is set (the operator writes it); after
// pointer swaps, the bits follow the buffer, not the argument position.
template <typename Op>
void run(const char *fn, vp_field **f, int n, unsigned up, unsigned down, Op op) {
    bool host = false, on_dev[3] = {false, false, false};
    for (int i = 0; i < n; ++i)
        if (f[i] && f[i]->data) {
            on_dev[i] = is_device(f[i]->data);
            host = host || !on_dev[i];
        }
    pfs_field pf[3];
    for (int i = 0; i < n; ++i)
        pf[i] = f[i] ? pfs_field{f[i]->x, f[i]->y, f[i]->z, f[i]->data} : pfs_field{0, 0, 0, nullptr};
    if (!host) {
        report(fn, op(pf));
        for (int i = 0; i < n; ++i)
            if (f[i]) f[i]->data = pf[i].data;
        return;
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k)
            if (f[i] && f[k] && f[i]->data && f[i]->data == f[k]->data) {
                std::fprintf(stderr, "pfs: %s: two arguments share one buffer\n", fn);
                return;
            }
    std::lock_guard<std::mutex> lock(g_mirror_mutex);
    float *hosts[3] = {nullptr, nullptr, nullptr};
    float *devs[3] = {nullptr, nullptr, nullptr};
    size_t sizes[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        if (!f[i] || !f[i]->data) continue;
        hosts[i] = f[i]->data;
        sizes[i] = nbytes(f[i]);
        if (on_dev[i]) {  // a device field among host ones: used in place
            devs[i] = hosts[i];
            continue;
        }
        devs[i] = mirror_of(hosts[i], sizes[i]);
        if (!devs[i]) {
            std::fprintf(stderr, "pfs: %s: device buffer allocation failed\n", fn);
            return;
        }
        pf[i].data = devs[i];
        if ((up >> i) & 1u) {
            if (hipMemcpy(devs[i], hosts[i], sizes[i], hipMemcpyHostToDevice) != hipSuccess) {
                std::fprintf(stderr, "pfs: %s: copy to the GPU failed\n", fn);
                return;
            }
        }
    }
    const int rc = op(pf);
    report(fn, rc);
    if (rc != PFS_OK) return;
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "pfs: %s: device synchronisation failed\n", fn);
        return;
    }
    // Argument i now holds device buffer pf[i].data: map it back to its host
    // buffer, and copy back the buffers the operator wrote.
    for (int i = 0; i < n; ++i) {
        if (!f[i] || !f[i]->data) continue;
        int src = -1;
        for (int k = 0; k < n; ++k)
            if (devs[k] == pf[i].data) src = k;
        if (src < 0) continue;  // not one of the mirrors: leave the struct as it was
        if (((down >> src) & 1u) && !on_dev[src]) {
            if (hipMemcpy(hosts[src], devs[src], sizes[src], hipMemcpyDeviceToHost) != hipSuccess) {
                std::fprintf(stderr, "pfs: %s: copy back to the host failed\n", fn);
                return;
            }
            down &= ~(1u << src);
        }
        f[i]->data = hosts[src];
    }
}

}  // namespace

void set_num_jacobi_iters(int iters) { g_iters.store(iters < 0 ? 0 : iters); }
int get_num_jacobi_iters(void) { return g_iters.load(); }

void advect(vp_field *vp, vp_field *vp_out, float dt) {
    vp_field *f[2] = {vp, vp_out};
    run("advect", f, 2, 0b11, 0b10, [&](pfs_field *p) { return pfs_advect(&p[0], &p[1], dt, nullptr); });
}

void advect_color(vp_field *image, vp_field *out, vp_field *vp, float dt) {
    vp_field *f[3] = {image, out, vp};
    run("advect_color", f, 3, 0b101, 0b010,
        [&](pfs_field *p) { return pfs_advect_color(&p[0], &p[1], &p[2], dt, nullptr); });
}

void diffuse(vp_field *vp, vp_field *vp_out, float viscosity, float dt) {
    vp_field *f[2] = {vp, vp_out};
    const int k = g_iters.load();
    run("diffuse", f, 2, 0b11, 0b11,
        [&](pfs_field *p) { return pfs_diffuse(&p[0], &p[1], viscosity, dt, k, nullptr); });
}

void addForces(vp_field *vp, float *forces) {
    (void)forces;
    if (!vp) std::fprintf(stderr, "pfs: addForces: vp is NULL\n");
}

void computePressure(vp_field *vp, vp_field *vp_out, float dt) {
    vp_field *f[2] = {vp, vp_out};
    const int k = g_iters.load();
    run("computePressure", f, 2, 0b11, 0b11,
        [&](pfs_field *p) { return pfs_compute_pressure(&p[0], &p[1], dt, k, nullptr); });
}

void subtractPressureGradient(vp_field *vp, vp_field *vp_out, float dt) {
    vp_field *f[2] = {vp, vp_out};
    run("subtractPressureGradient", f, 2, 0b11, 0b10, [&](pfs_field *p) {
        return pfs_subtract_pressure_gradient(&p[0], &p[1], dt, nullptr);
    });
}

void simulate_fluid_step(vp_field *vp, vp_field *tmp, float dt, float viscosity) {
    vp_field *f[2] = {vp, tmp};
    const int k = g_iters.load();
    run("simulate_fluid_step", f, 2, 0b11, 0b11, [&](pfs_field *p) {
        return pfs_simulate_fluid_step(&p[0], &p[1], dt, viscosity, k, nullptr);
    });
}

void advect_color_step(vp_field *image, vp_field *itmp, vp_field *vp, float dt) {
    vp_field *f[3] = {image, itmp, vp};
    run("advect_color_step", f, 3, 0b101, 0b010,
        [&](pfs_field *p) { return pfs_advect_color_step(&p[0], &p[1], &p[2], dt, nullptr); });
}

void simulate_fluid_step(float **vp, float **tmp, float dt, float viscosity, int vx, int vy,
                         int vz) {
    report("simulate_fluid_step",
           pfs_simulate_fluid_step_dev(vp, tmp, dt, viscosity, vx, vy, vz, g_iters.load(),
                                       nullptr));
}

void advect_color_step(float **image, float **itmp, float **vp, float dt, int ix, int iy, int iz,
                       int vx, int vy, int vz) {
    report("advect_color_step",
           pfs_advect_color_step_dev(image, itmp, vp, dt, ix, iy, iz, vx, vy, vz, nullptr));
}
