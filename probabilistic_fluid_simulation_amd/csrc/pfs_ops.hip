// Per-operator kernels: one launch per operator / per Jacobi sweep, one
// thread per cell, exactly the structure of the reference's src/fluid.cpp
// (and of the launch sequence of src/fluid.cu, but with fluid.cpp's results).
// These back the individual includes/fluid.hpp operators and the
// per-operator step path; the fused step lives in pfs_step.hip.
//
// All kernels address the interleaved [y][x][c] float field with 32-bit
// indices; the host rejects fields of 2^31 or more floats.

#include "pfs_device.h"
#include "pfs_internal.h"

namespace pfs {

constexpr int kBX = 64;  // one wavefront across x
constexpr int kBY = 4;

__global__ __launch_bounds__(kBX *kBY) void k_advect(const float *__restrict__ f,
                                                     float *__restrict__ o, int w, int h, int c,
                                                     float dt, float fw, float fh, float inv_w,
                                                     float inv_h, int pow2_w, int pow2_h) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int base = (j * w + i) * c;
    const float u = f[base], v = f[base + 1];
    const float xp = (float)i - div_n(dt * u, fw, inv_w, pow2_w);  // src/fluid.cpp:39
    const float yp = (float)j - div_n(dt * v, fh, inv_h, pow2_h);  // src/fluid.cpp:41
    const Trace t = trace(xp, yp, w, h, fw, fh);
    const int a00 = (t.j0 * w + t.i0) * c, a10 = (t.j0 * w + t.i1) * c;
    const int a01 = (t.j1 * w + t.i0) * c, a11 = (t.j1 * w + t.i1) * c;
#pragma unroll
    for (int k = 0; k < 2; ++k)
        o[base + k] = bilerp(t.sx, t.sy, f[a00 + k], f[a10 + k], f[a01 + k], f[a11 + k]);
}

constexpr int kColorRows = 4;  // dye pixels per thread, kBY rows apart

__global__ __launch_bounds__(kBX *kBY) void k_advect_color(ColorArgs a) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * (kBY * kColorRows) + threadIdx.y;
    advect_color_column<kColorRows>(a, i, j, kBY);
}

// One Jacobi smoothing sweep of channels 0,1 (src/fluid.cpp:154-185).
__global__ __launch_bounds__(kBX *kBY) void k_diffuse_sweep(const float *__restrict__ in,
                                                            float *__restrict__ out, int w, int h,
                                                            int c, float alpha, float beta) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int im = wrap_minus(i, w), ip = wrap_plus(i, w);
    const int jm = wrap_minus(j, h), jp = wrap_plus(j, h);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float left = in[(j * w + im) * c + k];
        const float right = in[(j * w + ip) * c + k];
        const float top = in[(jm * w + i) * c + k];
        const float bottom = in[(jp * w + i) * c + k];
        const float u_n = in[(j * w + i) * c + k];
        out[(j * w + i) * c + k] = diffuse_cell(left, right, top, bottom, u_n, alpha, beta);
    }
}

// Divergence into channel 3 of both fields (src/fluid.cpp:221-237).
__global__ __launch_bounds__(kBX *kBY) void k_divergence(float *__restrict__ in,
                                                         float *__restrict__ out, int w, int h,
                                                         int c, float gamma) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int im = wrap_minus(i, w), ip = wrap_plus(i, w);
    const int jm = wrap_minus(j, h), jp = wrap_plus(j, h);
    const float uR = in[(j * w + ip) * c + 0] - in[(j * w + im) * c + 0];
    const float vT = in[(jp * w + i) * c + 1] - in[(jm * w + i) * c + 1];
    const float b = gamma * (uR + vT);
    in[(j * w + i) * c + 3] = b;
    out[(j * w + i) * c + 3] = b;
}

// One pressure Jacobi sweep of channel 2 (src/fluid.cpp:240-257).
__global__ __launch_bounds__(kBX *kBY) void k_pressure_sweep(const float *__restrict__ in,
                                                             float *__restrict__ out, int w,
                                                             int h, int c) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int im = wrap_minus(i, w), ip = wrap_plus(i, w);
    const int jm = wrap_minus(j, h), jp = wrap_plus(j, h);
    const float pL = in[(j * w + im) * c + 2];
    const float pR = in[(j * w + ip) * c + 2];
    const float pT = in[(jm * w + i) * c + 2];
    const float pB = in[(jp * w + i) * c + 2];
    const float b = in[(j * w + i) * c + 3];
    out[(j * w + i) * c + 2] = pressure_cell(pL, pR, pT, pB, b);
}

// Velocity minus pressure gradient (src/fluid.cpp:273-295).
__global__ __launch_bounds__(kBX *kBY) void k_subtract_gradient(const float *__restrict__ in,
                                                                float *__restrict__ out, int w,
                                                                int h, int c, float dt) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int im = wrap_minus(i, w), ip = wrap_plus(i, w);
    const int jm = wrap_minus(j, h), jp = wrap_plus(j, h);
    const float pL = in[(j * w + im) * c + 2];
    const float pR = in[(j * w + ip) * c + 2];
    const float pT = in[(jm * w + i) * c + 2];
    const float pB = in[(jp * w + i) * c + 2];
    const float gradX = (pR - pL) * dt / 2.0f;
    const float gradY = (pB - pT) * dt / 2.0f;
    out[(j * w + i) * c + 0] = in[(j * w + i) * c + 0] - gradX;
    out[(j * w + i) * c + 1] = in[(j * w + i) * c + 1] - gradY;
}

// Debug dump of advect's back-trace keys (src/fluid.cpp:55-57).
__global__ __launch_bounds__(kBX *kBY) void k_advect_keys(const float *__restrict__ f,
                                                          int *__restrict__ keys, int w, int h,
                                                          int c, float dt, float fw, float fh,
                                                          float inv_w, float inv_h, int pow2_w,
                                                          int pow2_h) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int base = (j * w + i) * c;
    const float xp = (float)i - div_n(dt * f[base], fw, inv_w, pow2_w);
    const float yp = (float)j - div_n(dt * f[base + 1], fh, inv_h, pow2_h);
    const Trace t = trace(xp, yp, w, h, fw, fh);
    *reinterpret_cast<int4 *>(keys + 4 * (j * w + i)) = make_int4(t.i0, t.j0, t.i1, t.j1);
}

// Debug dump of advect_color's keys (src/fluid.cpp:89-90,112-113).
__global__ __launch_bounds__(kBX *kBY) void k_advect_color_keys(
    const float *__restrict__ vp, int *__restrict__ keys, int iw, int ih, int vw, int vh, int vc,
    float dt_viw, float dt_vih, float viw, float vih, float fiw, float fih, float inv_iw,
    float inv_ih, int pow2_iw, int pow2_ih) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= iw || j >= ih) return;
    const int vi = min((int)((float)i * viw), vw - 1);
    const int vj = min((int)((float)j * vih), vh - 1);
    const int vb = (vj * vw + vi) * vc;
    const float xp = (float)i - div_n(dt_viw * vp[vb], fiw, inv_iw, pow2_iw);
    const float yp = (float)j - div_n(dt_vih * vp[vb + 1], fih, inv_ih, pow2_ih);
    const Trace t = trace(xp, yp, iw, ih, fiw, fih);
    int *k = keys + 6 * (j * iw + i);
    k[0] = vi;
    k[1] = vj;
    k[2] = t.i0;
    k[3] = t.j0;
    k[4] = t.i1;
    k[5] = t.j1;
}

// Force injection (the step the reference leaves as a stub,
// src/fluid.cpp:198-208, called nowhere: src/fluid.cpp:302 comments it
// out): vp[cell][k] += dt * forces[cell][k] for channels k < fz, the
// product rounded, then the sum (no fused multiply-add).
__global__ __launch_bounds__(kBX *kBY) void k_add_forces(float *__restrict__ vp,
                                                         const float *__restrict__ f, int w, int h,
                                                         int c, int fz, float dt) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    if (i >= w || j >= h) return;
    const int cell = j * w + i;
    for (int k = 0; k < fz; ++k) vp[cell * c + k] = vp[cell * c + k] + dt * f[cell * fz + k];
}

namespace {
dim3 grid_for(int w, int h) { return dim3((w + kBX - 1) / kBX, (h + kBY - 1) / kBY); }
const dim3 kBlock(kBX, kBY);
}  // namespace

// ---- host launchers -------------------------------------------------------

int launch_advect(const float *f, float *o, int w, int h, int c, float dt, hipStream_t s) {
    const Pow2 pw(w), ph(h);
    hipLaunchKernelGGL(k_advect, grid_for(w, h), kBlock, 0, s, f, o, w, h, c, dt, (float)w,
                       (float)h, pw.inv, ph.inv, pw.is, ph.is);
    return check_launch("advect");
}

ColorArgs color_args(const float *img, float *out, const float *vp, int iw, int ih, int ic, int vw,
                     int vh, int vc, float dt) {
    const ColorConsts k(iw, ih, vw, vh, dt);
    ColorArgs a{};
    a.img = img;
    a.out = out;
    a.vp = vp;
    a.iw = iw;
    a.ih = ih;
    a.ic = ic;
    a.vw = vw;
    a.vh = vh;
    a.vc = vc;
    a.dt_viw = k.dt_viw;
    a.dt_vih = k.dt_vih;
    a.viw = k.viw;
    a.vih = k.vih;
    a.fiw = k.fiw;
    a.fih = k.fih;
    a.inv_iw = k.piw.inv;
    a.inv_ih = k.pih.inv;
    a.pow2_iw = k.piw.is;
    a.pow2_ih = k.pih.is;
    // float4 loads/stores need 4 channels and 16-byte aligned buffers.
    a.vec4 = ic == 4 && (uintptr_t)img % 16 == 0 && (uintptr_t)out % 16 == 0;
    a.blocks_x = (iw + kBX - 1) / kBX;
    return a;
}

int launch_advect_color(const float *img, float *out, const float *vp, int iw, int ih, int ic,
                        int vw, int vh, int vc, float dt, hipStream_t s) {
    const ColorArgs a = color_args(img, out, vp, iw, ih, ic, vw, vh, vc, dt);
    const dim3 grid((iw + kBX - 1) / kBX, (ih + kBY * kColorRows - 1) / (kBY * kColorRows));
    hipLaunchKernelGGL(k_advect_color, grid, kBlock, 0, s, a);
    return check_launch("advect_color");
}

int launch_diffuse_sweep(const float *in, float *out, int w, int h, int c, float alpha,
                         float beta, hipStream_t s) {
    hipLaunchKernelGGL(k_diffuse_sweep, grid_for(w, h), kBlock, 0, s, in, out, w, h, c, alpha,
                       beta);
    return check_launch("diffuse_sweep");
}

int launch_divergence(float *in, float *out, int w, int h, int c, float gamma, hipStream_t s) {
    hipLaunchKernelGGL(k_divergence, grid_for(w, h), kBlock, 0, s, in, out, w, h, c, gamma);
    return check_launch("divergence");
}

int launch_pressure_sweep(const float *in, float *out, int w, int h, int c, hipStream_t s) {
    hipLaunchKernelGGL(k_pressure_sweep, grid_for(w, h), kBlock, 0, s, in, out, w, h, c);
    return check_launch("pressure_sweep");
}

int launch_subtract_gradient(const float *in, float *out, int w, int h, int c, float dt,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_subtract_gradient, grid_for(w, h), kBlock, 0, s, in, out, w, h, c, dt);
    return check_launch("subtract_gradient");
}

int launch_add_forces(float *vp, const float *f, int w, int h, int c, int fz, float dt,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_add_forces, grid_for(w, h), kBlock, 0, s, vp, f, w, h, c, fz, dt);
    return check_launch("add_forces");
}

int launch_advect_keys(const float *f, int *keys, int w, int h, int c, float dt, hipStream_t s) {
    const Pow2 pw(w), ph(h);
    hipLaunchKernelGGL(k_advect_keys, grid_for(w, h), kBlock, 0, s, f, keys, w, h, c, dt,
                       (float)w, (float)h, pw.inv, ph.inv, pw.is, ph.is);
    return check_launch("advect_keys");
}

int launch_advect_color_keys(const float *vp, int *keys, int iw, int ih, int vw, int vh, int vc,
                             float dt, hipStream_t s) {
    const ColorConsts k(iw, ih, vw, vh, dt);
    hipLaunchKernelGGL(k_advect_color_keys, grid_for(iw, ih), kBlock, 0, s, vp, keys, iw, ih, vw,
                       vh, vc, k.dt_viw, k.dt_vih, k.viw, k.vih, k.fiw, k.fih, k.piw.inv,
                       k.pih.inv, k.piw.is, k.pih.is);
    return check_launch("advect_color_keys");
}

}  // namespace pfs
