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
// Each workgroup owns a TX x TY tile of cells and stages the tile plus a
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
// b = divergence of D^K.  Kernel D copies the pressure start to a workspace
// plane so that kernel P's halo reads never race with its own writes.

#include <cmath>

#include "pfs_device.h"
#include "pfs_internal.h"

namespace pfs {

namespace {

constexpr int kTile = 64;    // tile side
constexpr int kSegs = 6;     // row segments per staged column
constexpr int kMaxChunk = 8; // sweeps per launch

template <int H_>
struct Geo {
    static constexpr int H = H_;
    static constexpr int TX = kTile, TY = kTile;
    static constexpr int RX = TX + 2 * H, RY = TY + 2 * H;
    static constexpr int S = (RY + kSegs - 1) / kSegs;  // rows per thread
    static constexpr int NTH = RX * kSegs;              // threads owning cells
    static constexpr int NT = ((NTH + 63) / 64) * 64;   // whole wavefronts
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Remap so that each XCD owns a contiguous band of tiles and the
// halos its tiles share stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    if ((n & 7) != 0) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
}

__device__ __forceinline__ int wrap_index(int g, int n) {
    g %= n;
    return g < 0 ? g + n : g;
}

struct DiffuseArgs {
    const float *P;      // vp, [h][w][4]
    float *Q;            // tmp, [h][w][4]
    const float2 *dsrc;  // iterate plane read by a non-first chunk
    float2 *ddst;        // iterate plane written by a non-last chunk
    float *pw;           // pressure-start plane (last chunk)
    float *bpl;          // divergence plane (last chunk)
    int w, h, tiles_x, ntiles;
    float dt, fw, fh, inv_w, inv_h;
    int pow2_w, pow2_h;
    float alpha, beta, rcp_beta, gamma;
    int fast_div, first, k_even;
};

struct PressureArgs {
    float *P;
    float *Q;
    const float *psrc;  // iterate plane read by a non-first chunk
    float *pdst;        // iterate plane written by a non-last chunk
    const float *pw;
    const float *bpl;
    const float *image;  // dye, [h][w][4] (fused dye step only)
    float *itmp;
    int w, h, tiles_x, ntiles;
    float dt, fw, fh, inv_w, inv_h;
    int pow2_w, pow2_h;
    float dt_viw, dt_vih;  // dt / (field/image width ratio), as src/fluid.cpp:97-98
    int first;
};

// Per-thread view of the staged region.
template <class G>
struct Strip {
    bool active;
    int x, r0, xl, xr;
    __device__ explicit Strip(int tid) {
        active = tid < G::NTH;
        x = active ? tid % G::RX : 0;
        r0 = active ? (tid / G::RX) * G::S : 0;
        xl = x > 0 ? x - 1 : x;
        xr = x < G::RX - 1 ? x + 1 : x;
    }
    __device__ bool row_ok(int k) const { return active && r0 + k < G::RY; }
};

template <int H, bool LAST>
__global__ __launch_bounds__(Geo<H>::NT) void k_diffuse(DiffuseArgs a) {
    using G = Geo<H>;
    constexpr int S = G::S, RX = G::RX, RY = G::RY;
    constexpr int NS = H - (LAST ? 1 : 0);  // sweeps in this launch
    __shared__ float2 plane[2][RY * RX];
    __shared__ int wx[RX], wy[RY];

    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int x0 = (tile % a.tiles_x) * G::TX - H, y0 = (tile / a.tiles_x) * G::TY - H;
    for (int t = threadIdx.x; t < RX; t += G::NT) wx[t] = wrap_index(x0 + t, a.w);
    for (int t = threadIdx.x; t < RY; t += G::NT) wy[t] = wrap_index(y0 + t, a.h);
    __syncthreads();

    // Stage the iterate of every region cell in plane 0 (any thread may
    // produce any cell here), then pick up this thread's strip.
    {
        const float2 *P2 = reinterpret_cast<const float2 *>(a.P);  // (u, v) at 2*cell
        float2 *B0 = plane[0];
        for_rect<G::NT>(0, 0, RX, RY, [&](int rx, int ry) {
            const int gi = wx[rx], gj = wy[ry];
            float2 d;
            if (a.first) {
                // Advected velocity of the cell (src/fluid.cpp:36-67).
                const float2 uv = P2[2 * (gj * a.w + gi)];
                const float xp = (float)gi - div_n(a.dt * uv.x, a.fw, a.inv_w, a.pow2_w);
                const float yp = (float)gj - div_n(a.dt * uv.y, a.fh, a.inv_h, a.pow2_h);
                const Trace t = trace(xp, yp, a.w, a.h, a.fw, a.fh);
                const float2 f00 = P2[2 * (t.j0 * a.w + t.i0)];
                const float2 f10 = P2[2 * (t.j0 * a.w + t.i1)];
                const float2 f01 = P2[2 * (t.j1 * a.w + t.i0)];
                const float2 f11 = P2[2 * (t.j1 * a.w + t.i1)];
                d = make_float2(bilerp(t.sx, t.sy, f00.x, f10.x, f01.x, f11.x),
                                bilerp(t.sx, t.sy, f00.y, f10.y, f01.y, f11.y));
            } else {
                d = a.dsrc[gj * a.w + gi];
            }
            B0[ry * RX + rx] = d;
        });
    }
    __syncthreads();
    const Strip<G> st(threadIdx.x);
    float u[S], v[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const float2 d = st.row_ok(k) ? plane[0][(st.r0 + k) * RX + st.x] : make_float2(0.f, 0.f);
        u[k] = d.x;
        v[k] = d.y;
    }

    // Jacobi smoothing sweeps (src/fluid.cpp:154-185):
    // out = (a*L + a*R + a*T + a*B + 1*C) / beta, left to right.  Sweep s
    // publishes into plane s&1, so plane 0 is free again from sweep 2 on.
    // The division takes the 5-FMA exactly rounded path; numerators outside
    // its range are redone with IEEE division in one rare fix-up pass.
    float pu[S], pv[S];  // iterate before the last sweep (E for odd K)
    const bool fast = a.fast_div != 0;
    for (int s = 1; s <= NS; ++s) {
        float2 *B = plane[s & 1];
        float au[S], av[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            au[k] = a.alpha * u[k];
            av[k] = a.alpha * v[k];
            if (st.row_ok(k)) B[(st.r0 + k) * RX + st.x] = make_float2(au[k], av[k]);
        }
        __syncthreads();
        if (!st.active) continue;
        const float2 top = B[(st.r0 > 0 ? st.r0 - 1 : 0) * RX + st.x];
        const float2 bot = B[(st.r0 + S < RY ? st.r0 + S : RY - 1) * RX + st.x];
        float nu[S], nv[S];
        unsigned emin = 0xffu, emax = 0u;  // exponent range of the numerators
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (LAST && s == NS) {
                pu[k] = u[k];
                pv[k] = v[k];
            }
            const int row = (st.r0 + k < RY ? st.r0 + k : RY - 1) * RX;
            const float2 L = B[row + st.xl], R = B[row + st.xr];
            const float tu = k == 0 ? top.x : au[k - 1], tv = k == 0 ? top.y : av[k - 1];
            const float bu = k == S - 1 ? bot.x : au[k + 1];
            const float bv = k == S - 1 ? bot.y : av[k + 1];
            nu[k] = L.x + R.x + tu + bu + 1.0f * u[k];
            nv[k] = L.y + R.y + tv + bv + 1.0f * v[k];
            u[k] = div_fast(nu[k], a.beta, a.rcp_beta);
            v[k] = div_fast(nv[k], a.beta, a.rcp_beta);
            const unsigned eu = float_exp(nu[k]), ev = float_exp(nv[k]);
            emin = min(emin, min(eu, ev));
            emax = max(emax, max(eu, ev));
        }
        if (!fast || emin < kDivExpLo || emax > kDivExpHi) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                u[k] = div_fixed(nu[k], a.beta, a.rcp_beta, fast);
                v[k] = div_fixed(nv[k], a.beta, a.rcp_beta, fast);
            }
        }
    }

    if (!LAST) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int r = st.r0 + k;
            if (!st.row_ok(k) || r < H || r >= H + G::TY || st.x < H || st.x >= H + G::TX)
                continue;
            const int gx = x0 + st.x, gy = y0 + r;
            if (gx < a.w && gy < a.h) a.ddst[gy * a.w + gx] = make_float2(u[k], v[k]);
        }
        return;
    }

    // Divergence of D^K on the tile (src/fluid.cpp:221-237).
    float2 *B = plane[(NS + 1) & 1];
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (st.row_ok(k)) B[(st.r0 + k) * RX + st.x] = make_float2(u[k], v[k]);
    __syncthreads();
    float4 *Q4 = reinterpret_cast<float4 *>(a.Q);
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int r = st.r0 + k;
        if (!st.row_ok(k) || r < H || r >= H + G::TY || st.x < H || st.x >= H + G::TX) continue;
        const int gx = x0 + st.x, gy = y0 + r;
        if (gx >= a.w || gy >= a.h) continue;
        const int c = r * RX + st.x;
        const float uR = B[c + 1].x - B[c - 1].x;
        const float vT = B[c + RX].y - B[c - RX].y;
        const float b = a.gamma * (uR + vT);
        const int cell = gy * a.w + gx;
        const float eu = a.k_even ? u[k] : pu[k];
        const float ev = a.k_even ? v[k] : pv[k];
        const float4 q = Q4[cell];
        a.pw[cell] = a.k_even ? q.z : a.P[4 * cell + 2];
        a.bpl[cell] = b;
        Q4[cell] = make_float4(eu, ev, q.z, b);
    }
}

template <int H, bool LAST, bool DYE>
__global__ __launch_bounds__(Geo<H>::NT) void k_pressure(PressureArgs a) {
    using G = Geo<H>;
    constexpr int S = G::S, RX = G::RX, RY = G::RY;
    constexpr int NS = H - (LAST ? 1 : 0);
    __shared__ float plane[2][RY * RX];
    __shared__ int wx[RX], wy[RY];

    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int x0 = (tile % a.tiles_x) * G::TX - H, y0 = (tile / a.tiles_x) * G::TY - H;
    for (int t = threadIdx.x; t < RX; t += G::NT) wx[t] = wrap_index(x0 + t, a.w);
    for (int t = threadIdx.x; t < RY; t += G::NT) wy[t] = wrap_index(y0 + t, a.h);
    __syncthreads();
    const Strip<G> st(threadIdx.x);

    float p[S], rhs[S];
    {
        const float *src = a.first ? a.pw : a.psrc;
        const int gi = wx[st.x];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            p[k] = 0.0f;
            rhs[k] = 0.0f;
            if (!st.row_ok(k)) continue;
            const int g = wy[st.r0 + k] * a.w + gi;
            p[k] = src[g];
            rhs[k] = a.bpl[g];
        }
    }

    // Pressure Jacobi sweeps (src/fluid.cpp:240-257):
    // p' = (pL + pR + pT + pB + 1*b) / 4, left to right.
    float pp[S];  // p^(K-1)
    for (int s = 1; s <= NS; ++s) {
        float *B = plane[s & 1];
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.row_ok(k)) B[(st.r0 + k) * RX + st.x] = p[k];
        __syncthreads();
        if (st.active) {
            const float top = B[(st.r0 > 0 ? st.r0 - 1 : 0) * RX + st.x];
            const float bot = B[(st.r0 + S < RY ? st.r0 + S : RY - 1) * RX + st.x];
            float prev = top;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const float cur = p[k];
                if (LAST && s == NS) pp[k] = cur;
                if (st.r0 + k < RY) {
                    const int row = (st.r0 + k) * RX;
                    const float below = k == S - 1 ? bot : p[k + 1];
                    p[k] = pressure_cell(B[row + st.xl], B[row + st.xr], prev, below, rhs[k]);
                }
                prev = cur;
            }
        }
    }

    if (!LAST) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int r = st.r0 + k;
            if (!st.row_ok(k) || r < H || r >= H + G::TY || st.x < H || st.x >= H + G::TX)
                continue;
            const int gx = x0 + st.x, gy = y0 + r;
            if (gx < a.w && gy < a.h) a.pdst[gy * a.w + gx] = p[k];
        }
        return;
    }

    // Gradient subtraction on the tile (src/fluid.cpp:273-295).
    float *B = plane[(NS + 1) & 1];
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (st.row_ok(k)) B[(st.r0 + k) * RX + st.x] = p[k];
    __syncthreads();
    float4 *P4 = reinterpret_cast<float4 *>(a.P);
    float4 *Q4 = reinterpret_cast<float4 *>(a.Q);
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int r = st.r0 + k;
        if (!st.row_ok(k) || r < H || r >= H + G::TY || st.x < H || st.x >= H + G::TX) continue;
        const int gx = x0 + st.x, gy = y0 + r;
        if (gx >= a.w || gy >= a.h) continue;
        const int c = r * RX + st.x;
        const float gradX = (B[c + 1] - B[c - 1]) * a.dt / 2.0f;
        const float gradY = (B[c + RX] - B[c - RX]) * a.dt / 2.0f;
        const int cell = gy * a.w + gx;
        const float4 q = Q4[cell];  // (E.u, E.v, start, b)
        const float un = q.x - gradX, vn = q.y - gradY;
        P4[cell] = make_float4(un, vn, pp[k], rhs[k]);
        Q4[cell] = make_float4(q.x, q.y, p[k], rhs[k]);
        if (DYE) {
            // Dye advection of pixel (gx, gy) through the new velocity of
            // cell (gx, gy) (src/fluid.cpp:89-124, image size = field size).
            const float4 *I4 = reinterpret_cast<const float4 *>(a.image);
            const float xp = (float)gx - div_n(a.dt_viw * un, a.fw, a.inv_w, a.pow2_w);
            const float yp = (float)gy - div_n(a.dt_vih * vn, a.fh, a.inv_h, a.pow2_h);
            const Trace t = trace(xp, yp, a.w, a.h, a.fw, a.fh);
            const float4 f00 = I4[t.j0 * a.w + t.i0], f10 = I4[t.j0 * a.w + t.i1];
            const float4 f01 = I4[t.j1 * a.w + t.i0], f11 = I4[t.j1 * a.w + t.i1];
            reinterpret_cast<float4 *>(a.itmp)[cell] =
                make_float4(bilerp(t.sx, t.sy, f00.x, f10.x, f01.x, f11.x),
                            bilerp(t.sx, t.sy, f00.y, f10.y, f01.y, f11.y),
                            bilerp(t.sx, t.sy, f00.z, f10.z, f01.z, f11.z),
                            bilerp(t.sx, t.sy, f00.w, f10.w, f01.w, f11.w));
        }
    }
}

// ---- host dispatch over the compile-time halo ------------------------------
using DiffuseFn = void (*)(DiffuseArgs);
using PressureFn = void (*)(PressureArgs);

template <int H>
DiffuseFn diffuse_fn(bool last) {
    return last ? k_diffuse<H, true> : k_diffuse<H, false>;
}
template <int H>
PressureFn pressure_fn(bool last, bool dye) {
    if (!last) return k_pressure<H, false, false>;
    return dye ? k_pressure<H, true, true> : k_pressure<H, true, false>;
}

// Launch geometry by halo H = sweeps + (last ? 1 : 0), 1 <= H <= kMaxChunk+1.
int threads_for(int H) {
    switch (H) {
#define PFS_CASE(h) \
    case h:         \
        return Geo<h>::NT;
        PFS_CASE(2) PFS_CASE(3) PFS_CASE(4) PFS_CASE(5)
        PFS_CASE(6) PFS_CASE(7) PFS_CASE(8) PFS_CASE(9)
#undef PFS_CASE
    }
    return 0;
}

DiffuseFn diffuse_for(int H, bool last) {
    switch (H) {
        case 2: return diffuse_fn<2>(last);
        case 3: return diffuse_fn<3>(last);
        case 4: return diffuse_fn<4>(last);
        case 5: return diffuse_fn<5>(last);
        case 6: return diffuse_fn<6>(last);
        case 7: return diffuse_fn<7>(last);
        case 8: return diffuse_fn<8>(last);
        case 9: return diffuse_fn<9>(last);
    }
    return nullptr;
}

PressureFn pressure_for(int H, bool last, bool dye) {
    switch (H) {
        case 2: return pressure_fn<2>(last, dye);
        case 3: return pressure_fn<3>(last, dye);
        case 4: return pressure_fn<4>(last, dye);
        case 5: return pressure_fn<5>(last, dye);
        case 6: return pressure_fn<6>(last, dye);
        case 7: return pressure_fn<7>(last, dye);
        case 8: return pressure_fn<8>(last, dye);
        case 9: return pressure_fn<9>(last, dye);
    }
    return nullptr;
}

// Workspace layout (float offsets, N = w*h): pw [0, N), b [N, 2N),
// iterate planes d0 [2N, 4N), d1 [4N, 6N) (float2), p0 [6N, 7N), p1 [7N, 8N).
size_t ws_floats(int w, int h) { return 8 * (size_t)w * (size_t)h; }

}  // namespace

size_t fused_workspace_bytes(int w, int h) { return ws_floats(w, h) * sizeof(float); }

int fused_step(const StepArgs &s, float *ws, hipStream_t stream, hipEvent_t *ev) {
    const int w = s.w, h = s.h, K = s.iters;
    if (K < 1) return set_error(PFS_ERR_INVALID, "fused step needs iters >= 1");
    const size_t N = (size_t)w * h;
    const int tiles_x = (w + kTile - 1) / kTile, tiles_y = (h + kTile - 1) / kTile;
    const int ntiles = tiles_x * tiles_y;
    float *pw = ws, *bpl = ws + N;
    float2 *dpl[2] = {reinterpret_cast<float2 *>(ws + 2 * N),
                      reinterpret_cast<float2 *>(ws + 4 * N)};
    float *ppl[2] = {ws + 6 * N, ws + 7 * N};
    const Pow2 p2w(w), p2h(h);
    const float alpha = diffuse_alpha(s.visc, s.dt);
    const float beta = diffuse_beta(alpha);

    DiffuseArgs d{};
    d.P = s.P;
    d.Q = s.Q;
    d.pw = pw;
    d.bpl = bpl;
    d.w = w;
    d.h = h;
    d.tiles_x = tiles_x;
    d.ntiles = ntiles;
    d.dt = s.dt;
    d.fw = (float)w;
    d.fh = (float)h;
    d.inv_w = p2w.inv;
    d.inv_h = p2h.inv;
    d.pow2_w = p2w.is;
    d.pow2_h = p2h.is;
    d.alpha = alpha;
    d.beta = beta;
    d.rcp_beta = 1.0f / beta;
    d.fast_div = std::isfinite(beta) && std::fabs(beta) >= 0x1p-20f && std::fabs(beta) <= 0x1p20f;
    d.gamma = pressure_gamma(s.dt);
    d.k_even = (K % 2) == 0;

    if (ev) PFS_HIP_TRY(hipEventRecord(ev[0], stream));
    for (int done = 0, chunk = 0; done < K; ++chunk) {
        const int ns = (K - done) < kMaxChunk ? (K - done) : kMaxChunk;
        const bool last = done + ns == K;
        const int H = ns + (last ? 1 : 0);
        d.first = done == 0;
        d.dsrc = dpl[(chunk + 1) & 1];
        d.ddst = dpl[chunk & 1];
        hipLaunchKernelGGL(diffuse_for(H, last), dim3(ntiles), dim3(threads_for(H)), 0, stream,
                           d);
        PFS_TRY(check_launch("fused diffuse"));
        done += ns;
    }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[1], stream));

    PressureArgs p{};
    p.P = s.P;
    p.Q = s.Q;
    p.pw = pw;
    p.bpl = bpl;
    p.image = s.image;
    p.itmp = s.itmp;
    p.w = w;
    p.h = h;
    p.tiles_x = tiles_x;
    p.ntiles = ntiles;
    p.dt = s.dt;
    p.fw = (float)w;
    p.fh = (float)h;
    p.inv_w = p2w.inv;
    p.inv_h = p2h.inv;
    p.pow2_w = p2w.is;
    p.pow2_h = p2h.is;
    const bool dye = s.image != nullptr;
    if (dye) {
        const ColorConsts cc(w, h, w, h, s.dt);
        p.dt_viw = cc.dt_viw;
        p.dt_vih = cc.dt_vih;
    }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[2], stream));
    for (int done = 0, chunk = 0; done < K; ++chunk) {
        const int ns = (K - done) < kMaxChunk ? (K - done) : kMaxChunk;
        const bool last = done + ns == K;
        const int H = ns + (last ? 1 : 0);
        p.first = done == 0;
        p.psrc = ppl[(chunk + 1) & 1];
        p.pdst = ppl[chunk & 1];
        hipLaunchKernelGGL(pressure_for(H, last, dye && last), dim3(ntiles), dim3(threads_for(H)),
                           0, stream, p);
        PFS_TRY(check_launch("fused pressure"));
        done += ns;
    }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[3], stream));
    (void)tiles_y;
    return PFS_OK;
}

}  // namespace pfs
