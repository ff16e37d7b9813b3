// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "pfs.h"
#include "pfs_device.h"

namespace pfs {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);
void clear_error();

// Converts a pending launch error into PFS_ERR_HIP.
int check_launch(const char *what);

#define PFS_HIP_TRY(expr)                                                          \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return ::pfs::set_error(PFS_ERR_HIP, std::string(#expr) + ": " +       \
                                                     hipGetErrorString(e_));       \
    } while (0)

#define PFS_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != PFS_OK) return r_; \
    } while (0)

// 1/n and whether n is a power of two (then x/n == x*(1/n) exactly).
struct Pow2 {
    float inv;
    int is;
    explicit Pow2(int n) : inv(1.0f / (float)n), is(n > 0 && (n & (n - 1)) == 0) {}
};

// Host-side constants of advect_color, computed in float exactly as
// src/fluid.cpp:78-83,97-98 computes them.
ColorArgs color_args(const float *img, float *out, const float *vp, int iw, int ih, int ic, int vw,
                     int vh, int vc, float dt);

struct ColorConsts {
    float fiw, fih, viw, vih, dt_viw, dt_vih;
    Pow2 piw, pih;
    ColorConsts(int iw, int ih, int vw, int vh, float dt)
        : fiw((float)iw), fih((float)ih), viw((float)vw / (float)iw),
          vih((float)vh / (float)ih), dt_viw(dt / viw), dt_vih(dt / vih), piw(iw), pih(ih) {}
};

// diffuse's coefficients (src/fluid.cpp:144-145): beta is evaluated in double.
inline float diffuse_alpha(float visc, float dt) { return visc * dt; }
inline float diffuse_beta(float alpha) { return (float)(1.0 + 4.0 * (double)alpha); }
// computePressure's gamma (src/fluid.cpp:218): -1.0/dt in double, stored to float.
inline float pressure_gamma(float dt) { return (float)(-1.0 / (double)dt); }

// ---- per-operator launchers (pfs_ops.hip) --------------------------------
int launch_advect(const float *f, float *o, int w, int h, int c, float dt, hipStream_t s);
int launch_advect_color(const float *img, float *out, const float *vp, int iw, int ih, int ic,
                        int vw, int vh, int vc, float dt, hipStream_t s);
int launch_diffuse_sweep(const float *in, float *out, int w, int h, int c, float alpha,
                         float beta, hipStream_t s);
int launch_divergence(float *in, float *out, int w, int h, int c, float gamma, hipStream_t s);
int launch_pressure_sweep(const float *in, float *out, int w, int h, int c, hipStream_t s);
int launch_subtract_gradient(const float *in, float *out, int w, int h, int c, float dt,
                             hipStream_t s);
int launch_add_forces(float *vp, const float *f, int w, int h, int c, int fz, float dt,
                      hipStream_t s);
int launch_advect_keys(const float *f, int *keys, int w, int h, int c, float dt, hipStream_t s);
int launch_advect_color_keys(const float *vp, int *keys, int iw, int ih, int vw, int vh, int vc,
                             float dt, hipStream_t s);

// ---- fused step (pfs_step.hip) -------------------------------------------
// Tile sides the fused kernels are built for.  A small field has few tiles
// of the largest side, leaving most CUs idle; smaller tiles spread it over
// more workgroups, each with fewer rows per thread, at the price of a
// relatively wider halo.
constexpr int kTileSides[3] = {64, 32, 16};
inline bool valid_tile_side(int ts) { return ts == 64 || ts == 32 || ts == 16; }
int device_cus();  // CUs of the current device (256 if it does not answer)
// The tile side for a launch of this many tiles of side 64 and 32.
int tile_for_counts(long long tiles64, long long tiles32);
long long tile_count(int w, int h, int ts);

struct StepArgs {
    float *P;            // vp  (z == 4), device
    float *Q;            // tmp (z == 4), device
    const float *image;  // dye (w x h x 4) to advect in the same pass, or null
    float *itmp;         // its output buffer
    int w, h;
    float dt, visc;
    int iters;
    int tile;  // tile side of the fused kernels: 64, 32 or 16
};
// `ws` is a device workspace of fused_workspace_bytes(w, h).  Optional
// per-kernel event timing: if `ev` is non-null, events ev[2k] and ev[2k+1]
// are recorded around kernel group k (0 = diffuse, 1 = pressure).
// ev (profiling): events 0/1 around the diffusion kernels, 2/3 around the
// pressure kernels, each group launched `reps` times (both groups are
// idempotent: neither reads what it writes).
int fused_step(const StepArgs &a, float *ws, hipStream_t s, hipEvent_t *ev, int reps = 1);
size_t fused_workspace_bytes(int w, int h);
int threads_for(int H, int ts = 64);  // threads of the fused kernels for halo H, tile side ts

// Validation helpers.
int check_field(const pfs_field *f, const char *name, int min_channels);

}  // namespace pfs
