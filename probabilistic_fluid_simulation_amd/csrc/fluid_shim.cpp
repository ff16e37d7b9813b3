// The reference's includes/fluid.hpp API (include/fluid.hpp here) on top of
// the C ABI.  vp_field arguments may hold host memory — staged to the GPU
// and back on the null stream, returning with the results in host memory
// like the reference's CPU build — or device memory, enqueued asynchronously
// like the reference's CUDA build.  Pointer swaps of the reference
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

// Device copies of host fields for one call.  Maps each staged device
// pointer back to the host buffer it mirrors, so that after the call the
// caller's structs can be pointed at the host buffers in the same
// arrangement the device structs ended in.
struct Staging {
    struct Buf {
        float *host;
        float *dev;
        size_t bytes;
    };
    std::vector<Buf> bufs;
    int rc = PFS_OK;

    float *stage(float *host, size_t bytes, bool upload) {
        for (auto &b : bufs)
            if (b.host == host) return b.dev;
        void *d = nullptr;
        if (hipMalloc(&d, bytes) != hipSuccess) {
            rc = PFS_ERR_NOMEM;
            return nullptr;
        }
        if (upload && hipMemcpy(d, host, bytes, hipMemcpyHostToDevice) != hipSuccess)
            rc = PFS_ERR_HIP;
        bufs.push_back({host, static_cast<float *>(d), bytes});
        return static_cast<float *>(d);
    }
    float *host_of(float *dev) const {
        for (auto &b : bufs)
            if (b.dev == dev) return b.host;
        return dev;
    }
    void download() {
        for (auto &b : bufs)
            if (hipMemcpy(b.host, b.dev, b.bytes, hipMemcpyDeviceToHost) != hipSuccess)
                rc = PFS_ERR_HIP;
    }
    ~Staging() {
        for (auto &b : bufs) (void)hipFree(b.dev);
    }
};

// Runs `op` on pfs_field views of the n caller fields.  Host fields are
// staged to the device and every staged field is copied back afterwards.
template <typename Op>
void run(const char *fn, vp_field **f, int n, Op op) {
    bool host = false;
    for (int i = 0; i < n; ++i)
        if (f[i] && f[i]->data && !is_device(f[i]->data)) host = true;
    pfs_field pf[3];
    for (int i = 0; i < n; ++i)
        pf[i] = f[i] ? pfs_field{f[i]->x, f[i]->y, f[i]->z, f[i]->data} : pfs_field{0, 0, 0, nullptr};
    if (!host) {
        report(fn, op(pf));
        for (int i = 0; i < n; ++i)
            if (f[i]) f[i]->data = pf[i].data;
        return;
    }
    Staging st;
    for (int i = 0; i < n && st.rc == PFS_OK; ++i)
        if (f[i] && f[i]->data) pf[i].data = st.stage(f[i]->data, nbytes(f[i]), true);
    if (st.rc != PFS_OK) {
        std::fprintf(stderr, "pfs: %s: staging to the GPU failed\n", fn);
        return;
    }
    int rc = op(pf);
    report(fn, rc);
    if (rc != PFS_OK) return;
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "pfs: %s: device synchronisation failed\n", fn);
        return;
    }
    st.download();
    if (st.rc != PFS_OK) {
        std::fprintf(stderr, "pfs: %s: copy back to the host failed\n", fn);
        return;
    }
    for (int i = 0; i < n; ++i)
        if (f[i]) f[i]->data = st.host_of(pf[i].data);
}

}  // namespace

void set_num_jacobi_iters(int iters) { g_iters.store(iters < 0 ? 0 : iters); }
int get_num_jacobi_iters(void) { return g_iters.load(); }

void advect(vp_field *vp, vp_field *vp_out, float dt) {
    vp_field *f[2] = {vp, vp_out};
    run("advect", f, 2, [&](pfs_field *p) { return pfs_advect(&p[0], &p[1], dt, nullptr); });
}

void advect_color(vp_field *image, vp_field *out, vp_field *vp, float dt) {
    vp_field *f[3] = {image, out, vp};
    run("advect_color", f, 3,
        [&](pfs_field *p) { return pfs_advect_color(&p[0], &p[1], &p[2], dt, nullptr); });
}

void diffuse(vp_field *vp, vp_field *vp_out, float viscosity, float dt) {
    vp_field *f[2] = {vp, vp_out};
    const int k = g_iters.load();
    run("diffuse", f, 2,
        [&](pfs_field *p) { return pfs_diffuse(&p[0], &p[1], viscosity, dt, k, nullptr); });
}

void addForces(vp_field *vp, float *forces) {
    (void)forces;
    if (!vp) std::fprintf(stderr, "pfs: addForces: vp is NULL\n");
}

void computePressure(vp_field *vp, vp_field *vp_out, float dt) {
    vp_field *f[2] = {vp, vp_out};
    const int k = g_iters.load();
    run("computePressure", f, 2,
        [&](pfs_field *p) { return pfs_compute_pressure(&p[0], &p[1], dt, k, nullptr); });
}

void subtractPressureGradient(vp_field *vp, vp_field *vp_out, float dt) {
    vp_field *f[2] = {vp, vp_out};
    run("subtractPressureGradient", f, 2, [&](pfs_field *p) {
        return pfs_subtract_pressure_gradient(&p[0], &p[1], dt, nullptr);
    });
}

void simulate_fluid_step(vp_field *vp, vp_field *tmp, float dt, float viscosity) {
    vp_field *f[2] = {vp, tmp};
    const int k = g_iters.load();
    run("simulate_fluid_step", f, 2, [&](pfs_field *p) {
        return pfs_simulate_fluid_step(&p[0], &p[1], dt, viscosity, k, nullptr);
    });
}

void advect_color_step(vp_field *image, vp_field *itmp, vp_field *vp, float dt) {
    vp_field *f[3] = {image, itmp, vp};
    run("advect_color_step", f, 3,
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
