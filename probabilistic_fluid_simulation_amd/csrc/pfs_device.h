// Device helpers shared by the fluid-step kernels (gfx950, wave64).
//
// Floating-point contract: every expression below keeps the operation order
// of the reference's CPU solver (src/fluid.cpp); the library is compiled
// with -ffp-contract=off (no fused multiply-add), IEEE division and
// denormals preserved, so each kernel reproduces fluid.cpp bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace pfs {

// IEEE fmod in the field's precision (fmodf for float, as the reference's
// C++ overload resolution picks for float arguments, src/fluid.cpp:48).
__device__ __forceinline__ float fmod_t(float x, float y) { return fmodf(x, y); }
__device__ __forceinline__ double fmod_t(double x, double y) { return fmod(x, y); }

// fmod(fmod(x, n) + n, n) for an integer-valued n > 0 (src/fluid.cpp:48-49,
// 105-106), exactly.  When |x| < n, fmod(x, n) == x; x + n is rounded once
// to s in [0, 2n] and fmod(s, n) is s, s - n (exact by Sterbenz' lemma for
// s in [n, 2n)) or 0 for s == 2n.  Other inputs take the exact libm fmod.
template <typename T>
__device__ __forceinline__ T wrap_coord(T x, T n) {
    if (fabs(x) < n) {
        T s = x + n;
        s = (s >= n) ? s - n : s;
        return (s >= n) ? s - n : s;
    }
    return fmod_t(fmod_t(x, n) + n, n);
}

// x / n, correctly rounded.  For n a power of two the product with 1/n is the
// same correctly rounded value of the same real number.
template <typename T>
__device__ __forceinline__ T div_n(T x, T n, T inv_n, bool n_pow2) {
    return n_pow2 ? x * inv_n : x / n;
}

// Bilinear blend in the reference's exact order (src/fluid.cpp:62-66).
template <typename T>
__device__ __forceinline__ T bilerp(T sx, T sy, T f00, T f10, T f01, T f11) {
    return (1 - sx) * (1 - sy) * f00 + sx * (1 - sy) * f10 + (1 - sx) * sy * f01 +
           sx * sy * f11;
}

// Cell keys of one back-trace (src/fluid.cpp:48-58).
template <typename T>
struct TraceT {
    int i0, j0, i1, j1;
    T sx, sy;
};
using Trace = TraceT<float>;

// Wraps (xp, yp) into [0, w) x [0, h) and returns the corner keys and weights.
// The clamp of i0/j0 is a no-op for every finite input (the wrapped
// coordinate lies in [0, n)); it only keeps a NaN/Inf velocity, for which
// the reference indexes memory out of bounds, inside the field.
template <typename T>
__device__ __forceinline__ TraceT<T> trace(T xp, T yp, int w, int h, T fw, T fh) {
    xp = wrap_coord(xp, fw);
    yp = wrap_coord(yp, fh);
    TraceT<T> t;
    t.i0 = min(max((int)xp, 0), w - 1);
    t.j0 = min(max((int)yp, 0), h - 1);
    t.i1 = (t.i0 + 1 == w) ? 0 : t.i0 + 1;  // (i0 + 1) % w for i0 in [0, w)
    t.j1 = (t.j0 + 1 == h) ? 0 : t.j0 + 1;
    t.sx = xp - (T)t.i0;
    t.sy = yp - (T)t.j0;
    return t;
}

// Periodic neighbours, equal to ((i-1) % n + n) % n and (i+1) % n for i in [0, n).
__device__ __forceinline__ int wrap_minus(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrap_plus(int i, int n) { return i + 1 == n ? 0 : i + 1; }

// Diffusion sweep of one channel (src/fluid.cpp:19-21,168-182).
template <typename T>
__device__ __forceinline__ T diffuse_cell(T l, T r, T t, T b, T c, T alpha, T beta) {
    return (alpha * l + alpha * r + alpha * t + alpha * b + (T)1.0f * c) / beta;
}

// x / beta correctly rounded, for a divisor fixed per launch, with y =
// RN(1/beta) precomputed.  q1 = q0 + (x - q0*beta)*y is within one ulp of
// x/beta (faithful), so by Markstein's theorem (Markstein 1990; Cornea,
// Harrison & Tang 2002, final-step theorem) q1 + (x - q1*beta)*y rounded
// once is RN(x/beta).  The remainders are exact with FMA as long as nothing
// under- or overflows, which the host guarantees by allowing `fast` only for
// 2^-20 <= beta <= 2^20 and the range test below keeps for x; every other x
// (zero, tiny, huge, non-finite) takes IEEE division.
__device__ __forceinline__ bool div_fast_ok(float x) {
    const float ax = fabsf(x);
    return ax >= 0x1p-96f && ax < 0x1p96f;  // false for 0, tiny, huge, Inf, NaN
}
// Biased exponent of x: div_fast_ok(x) <=> kDivExpLo <= float_exp(x) <= kDivExpHi.
// Lets a loop track min/max exponents branch-free and test the range once.
constexpr unsigned kDivExpLo = 127 - 96, kDivExpHi = 127 + 95;
__device__ __forceinline__ unsigned float_exp(float x) {
    return (__float_as_uint(x) >> 23) & 0xffu;
}
// One Markstein step: q + (x - q*beta)*y, each term rounded once.
__device__ __forceinline__ float div_step(float q, float x, float beta, float y) {
    return fmaf(fmaf(-q, beta, x), y, q);
}
__device__ __forceinline__ float div_fast(float x, float beta, float y) {
    return div_step(div_step(x * y, x, beta, y), x, beta, y);
}
// When the host has checked |y*beta - 1| <= 2^-26 (exact in double), then
// |x*y - x/beta| < ulp(x/beta)/4 for every x, so q0 = RN(x*y) is already
// faithful and one Markstein step rounds correctly: 3 operations instead
// of 5 (same range conditions on x).
__device__ __forceinline__ float div_fast3(float x, float beta, float y) {
    return div_step(x * y, x, beta, y);
}
__device__ __forceinline__ float div_fixed(float x, float beta, float y, bool fast) {
    return (fast && div_fast_ok(x)) ? div_fast(x, beta, y) : x / beta;
}

// Pressure sweep (src/fluid.cpp:249-255): jacobi(pL, pR, pT, pB, 1, 4, b).
template <typename T>
__device__ __forceinline__ T pressure_cell(T l, T r, T t, T b, T rhs) {
    return (l + r + t + b + (T)1.0f * rhs) / (T)4.0f;
}

// Constants of advect_color (src/fluid.cpp:78-83,97-98), computed on the
// host in float exactly as the reference computes them.
struct ColorArgs {
    const float *img;
    float *out;
    const float *vp;
    int iw, ih, ic, vw, vh, vc;
    float dt_viw, dt_vih, viw, vih, fiw, fih, inv_iw, inv_ih;
    int pow2_iw, pow2_ih, vec4;
    int blocks_x, block_base;  // 64x4-pixel blocks per row; first block in a batch
};

// One pixel of advect_color (src/fluid.cpp:86-125).
__device__ __forceinline__ void advect_color_pixel(const ColorArgs &a, int i, int j) {
    // Nearest velocity cell (src/fluid.cpp:89-90); clamped only against the
    // float rounding of i*viw reaching vw, which the reference would read
    // out of bounds.
    const int vi = min((int)((float)i * a.viw), a.vw - 1);
    const int vj = min((int)((float)j * a.vih), a.vh - 1);
    const int vb = (vj * a.vw + vi) * a.vc;
    const float xp = (float)i - div_n(a.dt_viw * a.vp[vb], a.fiw, a.inv_iw, a.pow2_iw);      // :97
    const float yp = (float)j - div_n(a.dt_vih * a.vp[vb + 1], a.fih, a.inv_ih, a.pow2_ih);  // :98
    const Trace t = trace(xp, yp, a.iw, a.ih, a.fiw, a.fih);
    const int ic = a.ic;
    const int base = (j * a.iw + i) * ic;
    const int a00 = (t.j0 * a.iw + t.i0) * ic, a10 = (t.j0 * a.iw + t.i1) * ic;
    const int a01 = (t.j1 * a.iw + t.i0) * ic, a11 = (t.j1 * a.iw + t.i1) * ic;
    if (a.vec4) {
        const float4 f00 = *reinterpret_cast<const float4 *>(a.img + a00);
        const float4 f10 = *reinterpret_cast<const float4 *>(a.img + a10);
        const float4 f01 = *reinterpret_cast<const float4 *>(a.img + a01);
        const float4 f11 = *reinterpret_cast<const float4 *>(a.img + a11);
        float4 r;
        r.x = bilerp(t.sx, t.sy, f00.x, f10.x, f01.x, f11.x);
        r.y = bilerp(t.sx, t.sy, f00.y, f10.y, f01.y, f11.y);
        r.z = bilerp(t.sx, t.sy, f00.z, f10.z, f01.z, f11.z);
        r.w = bilerp(t.sx, t.sy, f00.w, f10.w, f01.w, f11.w);
        *reinterpret_cast<float4 *>(a.out + base) = r;
    } else {
        for (int k = 0; k < ic; ++k)
            a.out[base + k] =
                bilerp(t.sx, t.sy, a.img[a00 + k], a.img[a10 + k], a.img[a01 + k], a.img[a11 + k]);
    }
}

// R pixels of column i, rows j0, j0 + dj, ... (those inside the image): the
// operations of advect_color_pixel for each, with every gather of the R
// pixels issued before any of their results is computed, so that the
// gathers of different pixels overlap.  Rows past the image load the last
// row (in bounds) and store nothing.
template <int R>
__device__ __forceinline__ void advect_color_column(const ColorArgs &a, int i, int j0, int dj) {
    if (i >= a.iw) return;
    if (!a.vec4) {
        for (int r = 0; r < R; ++r)
            if (j0 + r * dj < a.ih) advect_color_pixel(a, i, j0 + r * dj);
        return;
    }
    const int vi = min((int)((float)i * a.viw), a.vw - 1);
    float u[R], v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = min(j0 + r * dj, a.ih - 1);
        const int vj = min((int)((float)j * a.vih), a.vh - 1);
        const int vb = (vj * a.vw + vi) * a.vc;
        u[r] = a.vp[vb];
        v[r] = a.vp[vb + 1];
    }
    Trace t[R];
    float4 f[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = min(j0 + r * dj, a.ih - 1);
        const float xp = (float)i - div_n(a.dt_viw * u[r], a.fiw, a.inv_iw, a.pow2_iw);  // :97
        const float yp = (float)j - div_n(a.dt_vih * v[r], a.fih, a.inv_ih, a.pow2_ih);  // :98
        t[r] = trace(xp, yp, a.iw, a.ih, a.fiw, a.fih);
        const float4 *img = reinterpret_cast<const float4 *>(a.img);
        f[r][0] = img[t[r].j0 * a.iw + t[r].i0];
        f[r][1] = img[t[r].j0 * a.iw + t[r].i1];
        f[r][2] = img[t[r].j1 * a.iw + t[r].i0];
        f[r][3] = img[t[r].j1 * a.iw + t[r].i1];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r * dj;
        if (j >= a.ih) continue;
        const Trace &q = t[r];
        float4 o;
        o.x = bilerp(q.sx, q.sy, f[r][0].x, f[r][1].x, f[r][2].x, f[r][3].x);
        o.y = bilerp(q.sx, q.sy, f[r][0].y, f[r][1].y, f[r][2].y, f[r][3].y);
        o.z = bilerp(q.sx, q.sy, f[r][0].z, f[r][1].z, f[r][2].z, f[r][3].z);
        o.w = bilerp(q.sx, q.sy, f[r][0].w, f[r][1].w, f[r][2].w, f[r][3].w);
        reinterpret_cast<float4 *>(a.out)[j * a.iw + i] = o;
    }
}

// Visits every cell of the rectangle [x0, x0+rw) x [y0, y0+rh) once across
// the NT threads of a block, row-major, with an incremental (x, y) update
// instead of a division per cell.
template <int NT, typename F>
__device__ __forceinline__ void for_rect(int x0, int y0, int rw, int rh, F &&f) {
    const int n = rw * rh;
    int idx = (int)threadIdx.x;
    if (rw <= 0 || idx >= n) return;
    int ry = idx / rw;
    int rx = idx - ry * rw;
    const int dq = NT / rw, dr = NT - dq * rw;
    for (; idx < n; idx += NT) {
        f(x0 + rx, y0 + ry);
        rx += dr;
        ry += dq;
        if (rx >= rw) {
            rx -= rw;
            ry += 1;
        }
    }
}

}  // namespace pfs
