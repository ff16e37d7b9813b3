// Device side of the fused, temporally blocked fluid step (gfx950): the
// per-tile bodies shared by the single-scene kernels (pfs_step.hip) and the
// batched kernels (pfs_batch.hip).  The scheme is described in
// pfs_step.hip.
#pragma once

#include "pfs_device.h"

namespace pfs {

constexpr int kTile = 64;    // the largest tile side (the default)
constexpr int kSegs = 12;    // row segments per staged column of a 64-wide tile
constexpr int kMaxChunk = 8; // sweeps per launch

template <int H_, int TS_ = kTile, int NTMAX = 1024>
struct Geo {
    static constexpr int H = H_;
    static constexpr int TX = TS_, TY = TS_;
    static constexpr int RX = TX + 2 * H, RY = TY + 2 * H;
    // Rows per thread: a 64-wide tile keeps 12 segments per column; smaller
    // tiles take as many segments as 1024 threads allow, rows split evenly.
    // NTMAX < 1024: workgroups of at most NTMAX threads (longer strips), so
    // that two fit a CU.
    static constexpr int MAXSEG = (TS_ == kTile && NTMAX == 1024) ? kSegs : NTMAX / RX;
    static constexpr int S = (RY + MAXSEG - 1) / MAXSEG;
    static constexpr int SEGS = (TS_ == kTile && NTMAX == 1024) ? kSegs : (RY + S - 1) / S;
    static constexpr int RYP = SEGS * S;                // rows of an LDS plane (>= RY)
    static constexpr int NTH = RX * SEGS;               // threads owning cells
    static constexpr int NT = ((NTH + 63) / 64) * 64;   // whole wavefronts
    // Row pitch of an FP32 LDS plane: the smallest P >= RX with S*P = RX
    // (mod 64).  Thread t = seg*RX + x addresses row seg*S + k at
    // (seg*S + k)*P + x, which is t + k*P (mod 64): a wavefront whose strips
    // start in two segments still addresses 64 consecutive banks.  (The
    // pressure planes use it; for the diffusion's float2 planes it removed
    // conflicts without changing the kernel's time, so they keep pitch RX.)
    static constexpr int pitch() {
        for (int p = RX; p < RX + 64; ++p)
            if ((S * p - RX) % 64 == 0) return p;
        return RX;
    }
    static constexpr int P = pitch();
    static_assert(NTH <= NTMAX && RYP >= RY, "tile geometry");
};

// Workgroups of `threads` threads and `lds` bytes of LDS that one CU holds at
// once (160 KiB of LDS, 32 wavefronts).
constexpr int wgs_per_cu(int lds, int threads) {
    const int by_lds = 163840 / lds, by_waves = 32 / ((threads + 63) / 64);
    return by_lds < by_waves ? by_lds : by_waves;
}

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
    const float *Q;      // tmp, [h][w][4] (read only: its pressure, for even K)
    const float2 *dsrc;  // iterate plane read by a non-first chunk
    float2 *ddst;        // iterate plane written by a non-last chunk; E (last chunk)
    float *pw;           // pressure-start plane (last chunk)
    float *bpl;          // divergence plane (last chunk)
    int w, h, tiles_x, ntiles;
    float dt, fw, fh, inv_w, inv_h;
    int pow2_w, pow2_h;
    float alpha, beta, rcp_beta, gamma;
    int fast_div;  // 0: IEEE division; 1: 5-op exact division; 2: 3-op exact division
    int first, k_even;
    int tile_base;  // first tile of this scene in a batched launch
};

struct PressureArgs {
    float *P;
    float *Q;
    const float *psrc;  // iterate plane read by a non-first chunk
    float *pdst;        // iterate plane written by a non-last chunk
    const float *pw;
    const float *bpl;
    const float2 *E;     // the diffused velocity kept for the output (kernel D's last chunk)
    const float *image;  // dye, [h][w][4], advected in the same pass; null: no dye
    float *itmp;
    int w, h, tiles_x, ntiles;
    float dt, fw, fh, inv_w, inv_h;
    int pow2_w, pow2_h;
    float dt_viw, dt_vih;  // dt / (field/image width ratio), as src/fluid.cpp:97-98
    int first;
    int tile_base;  // first tile of this scene in a batched launch
};

// Thread (x, seg) of a tile: column x of the staged region, rows
// [r0, r0 + S), and the global cells those map to (periodic wrap).
template <class G>
struct Strip {
    bool active;
    int x, r0, xl, xr;
    int x0, y0;  // global coordinates of region cell (0, 0), unwrapped
    int gi;      // global column of x
    int gj[G::S];
    __device__ Strip(int tid, int tile, int tiles_x, int w, int h) {
        active = tid < G::NTH;
        x = active ? tid % G::RX : 0;
        r0 = active ? (tid / G::RX) * G::S : 0;
        xl = x > 0 ? x - 1 : x;
        xr = x < G::RX - 1 ? x + 1 : x;
        x0 = (tile % tiles_x) * G::TX - G::H;
        y0 = (tile / tiles_x) * G::TY - G::H;
        gi = wrap_index(x0 + x, w);
        int g = wrap_index(y0 + r0, h);
#pragma unroll
        for (int k = 0; k < G::S; ++k) {
            gj[k] = g;
            g = (g + 1 == h) ? 0 : g + 1;
        }
    }
    __device__ bool row_ok(int k) const { return active && r0 + k < G::RY; }
    // Row r0+k of column x is a tile cell inside the field (not halo or
    // past the field's edge); then its global cell is (x0+x, y0+r0+k).
    __device__ bool mine(int k, int w, int h) const {
        const int r = r0 + k;
        return active && r >= G::H && r < G::H + G::TY && x >= G::H && x < G::H + G::TX &&
               x0 + x < w && y0 + r < h;
    }
    __device__ int cell(int k, int w) const { return (y0 + r0 + k) * w + x0 + x; }
};

// Error-free transformation of a sum (Knuth's TwoSum): s = RN(a + b) and
// a + b = s + e exactly.
__device__ __forceinline__ float two_sum_err(float a, float b, float s) {
    const float bb = s - a;
    return (a - (s - bb)) + (b - bb);
}

// alpha*(L + R + T + B) + C, the numerator of the diffusion sweep
// (src/fluid.cpp:19-21,175-182: alpha*L + alpha*R + alpha*T + alpha*B + 1*C,
// the same real number), evaluated compensated: the neighbour sum by Kahan's
// compensated summation (running sum s, correction c, the sum being s - c;
// Kahan 1965), its product by alpha with the exact FMA error (TwoProduct)
// less alpha times c, the centre added by TwoSum (the two terms are not
// ordered by magnitude), and one final rounding.
__device__ __forceinline__ float diffuse_num_comp(float l, float r, float t, float b, float c,
                                                  float alpha) {
    float s = l + r;
    float k = (s - l) - r;
    float y = t - k, u = s + y;
    k = (u - s) - y;
    s = u;
    y = b - k;
    u = s + y;
    k = (u - s) - y;
    s = u;
    const float p = alpha * s;
    float ep = fmaf(alpha, s, -p);
    ep = fmaf(-alpha, k, ep);
    const float q = p + c;
    return q + (two_sum_err(p, c, q) + ep);
}

// L + R + T + B + rhs of the pressure sweep (src/fluid.cpp:249-255) by
// Kahan's compensated summation: s the running FP32 sum, k its correction,
// s - k the sum in about twice FP32 precision.
__device__ __forceinline__ void pressure_sum_comp(float l, float r, float t, float b, float rhs,
                                                  float &s, float &k) {
    s = l + r;
    k = (s - l) - r;
    float y = t - k, u = s + y;
    k = (u - s) - y;
    s = u;
    y = b - k;
    u = s + y;
    k = (u - s) - y;
    s = u;
    y = rhs - k;
    u = s + y;
    k = (u - s) - y;
    s = u;
}

// Region coordinate of global index g (0 <= g < n) on an axis whose staged
// region starts at global coordinate o (periodic), or `lim` when g + 1 is
// not staged right after it.
__device__ __forceinline__ int staged(int g, int o, int n, int lim) {
    int r = g - o;  // in (-n, n + H) for o in [-H, n)
    if (r < 0) r += n;
    if (r >= n) r -= n;
    return r < lim - 1 ? r : lim;
}

// COMP (the precision form): the numerators are compensated
// (diffuse_num_comp); the strip publishes its raw values instead of
// alpha-scaled ones.  COMP = false is the reference's arithmetic.
template <int H, bool LAST, int TS = kTile, bool COMP = false>
__device__ __forceinline__ void diffuse_tile(const DiffuseArgs &a, int tile) {
    using G = Geo<H, TS>;
    constexpr int S = G::S, RX = G::RX, RY = G::RY, PL = G::RYP * RX;
    constexpr int NS = H - (LAST ? 1 : 0);  // sweeps in this launch
    __shared__ float2 plane[2 * PL];

    const Strip<G> st(threadIdx.x, tile, a.tiles_x, a.w, a.h);
    const float2 *__restrict__ P2 = reinterpret_cast<const float2 *>(a.P);  // (u,v) at 2*cell

    // The staged region, loaded at entry: the advection needs all of it.
    float2 c[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int g = st.row_ok(k) ? st.gj[k] * a.w + st.gi : 0;
        c[k] = a.first ? P2[2 * g] : a.dsrc[g];
    }
    float u[S], v[S];
    if (a.first) {
        // Advect every strip cell (src/fluid.cpp:36-67) with its corners
        // from the staged velocities in LDS plane 1 — or from memory when a
        // back-trace leaves the staged region.
        float2 *V = plane + PL;
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.active) V[(st.r0 + k) * RX + st.x] = c[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const float xp = (float)st.gi - div_n(a.dt * c[k].x, a.fw, a.inv_w, a.pow2_w);
            const float yp = (float)st.gj[k] - div_n(a.dt * c[k].y, a.fh, a.inv_h, a.pow2_h);
            const Trace t = trace(xp, yp, a.w, a.h, a.fw, a.fh);
            const int rx = staged(t.i0, st.x0, a.w, RX), ry = staged(t.j0, st.y0, a.h, RY);
            float2 f00, f10, f01, f11;
            if (rx < RX && ry < RY) {
                const int o = ry * RX + rx;
                f00 = V[o];
                f10 = V[o + 1];
                f01 = V[o + RX];
                f11 = V[o + RX + 1];
            } else {
                f00 = P2[2 * (t.j0 * a.w + t.i0)];
                f10 = P2[2 * (t.j0 * a.w + t.i1)];
                f01 = P2[2 * (t.j1 * a.w + t.i0)];
                f11 = P2[2 * (t.j1 * a.w + t.i1)];
            }
            u[k] = bilerp(t.sx, t.sy, f00.x, f10.x, f01.x, f11.x);
            v[k] = bilerp(t.sx, t.sy, f00.y, f10.y, f01.y, f11.y);
        }
    } else {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            u[k] = c[k].x;
            v[k] = c[k].y;
        }
    }

    // The tile cells' pressure start (last chunk), needed only at the end:
    // issued after the advection, so that these loads (a whole 16-byte cell
    // for 4 bytes when they come from tmp) stream in under the sweeps instead
    // of competing with the staged region at entry.
    float warm[S];
    if (LAST) {
        const float *__restrict__ W = a.k_even ? a.Q : a.P;
#pragma unroll
        for (int k = 0; k < S; ++k)  // other rows load cell 0, unused
            warm[k] = W[4 * (st.mine(k, a.w, a.h) ? st.cell(k, a.w) : 0) + 2];
    }

    // Jacobi smoothing sweeps (src/fluid.cpp:154-185):
    // out = (a*L + a*R + a*T + a*B + 1*C) / beta, left to right.  Sweep s
    // publishes alpha-scaled cells into plane (s+1)&1: sweep 1 uses plane 0
    // while plane 1 may still be read by the advection above.  The division
    // takes an exactly rounded FMA path (3 or 5 operations); if any numerator
    // of the strip is outside that path's range (zero, tiny, huge, Inf) the
    // strip divides per element instead (IEEE division where needed).  Rows
    // past RY are padding: computed, never used.
    float pu[S], pv[S];  // iterate before the last sweep (E for odd K)
    const bool fast = a.fast_div != 0;
    for (int s = 1; s <= NS; ++s) {
        float2 *B = plane + ((s + 1) & 1) * PL;
        float au[S], av[S];  // what the strip publishes: alpha-scaled, or raw (COMP)
#pragma unroll
        for (int k = 0; k < S; ++k) {
            au[k] = COMP ? u[k] : a.alpha * u[k];
            av[k] = COMP ? v[k] : a.alpha * v[k];
            if (st.active) B[(st.r0 + k) * RX + st.x] = make_float2(au[k], av[k]);
        }
        __syncthreads();
        if (!st.active) continue;
        const float2 top = B[(st.r0 > 0 ? st.r0 - 1 : 0) * RX + st.x];
        const float2 bot = B[(st.r0 + S < RY ? st.r0 + S : RY - 1) * RX + st.x];
        const float2 *Bl = B + st.r0 * RX + st.xl, *Br = B + st.r0 * RX + st.xr;
        float nu[S], nv[S];
        float lo = 0x1p127f, hi = 0.0f;  // magnitude range of the numerators
        const bool short_div = a.fast_div == 2;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (LAST && s == NS) {
                pu[k] = u[k];
                pv[k] = v[k];
            }
            const float2 L = Bl[k * RX], R = Br[k * RX];
            const float tu = k == 0 ? top.x : au[k - 1], tv = k == 0 ? top.y : av[k - 1];
            const float bu = k == S - 1 ? bot.x : au[k + 1];
            const float bv = k == S - 1 ? bot.y : av[k + 1];
            if constexpr (COMP) {
                nu[k] = diffuse_num_comp(L.x, R.x, tu, bu, u[k], a.alpha);
                nv[k] = diffuse_num_comp(L.y, R.y, tv, bv, v[k], a.alpha);
            } else {
                nu[k] = L.x + R.x + tu + bu + 1.0f * u[k];
                nv[k] = L.y + R.y + tv + bv + 1.0f * v[k];
            }
            lo = fminf(lo, fminf(fabsf(nu[k]), fabsf(nv[k])));
            hi = fmaxf(hi, fmaxf(fabsf(nu[k]), fabsf(nv[k])));
        }
        // The range is tested before dividing, in the block that made the
        // numerators (the compiler then keeps min3/max3 on them).
        // div_fast is div_fast3 followed by one more Markstein step: the
        // step is applied in place, so both forms leave the quotients in the
        // same registers (no copies on the 3-operation path).
        if (fast && lo >= 0x1p-96f && hi < 0x1p96f) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                u[k] = div_fast3(nu[k], a.beta, a.rcp_beta);
                v[k] = div_fast3(nv[k], a.beta, a.rcp_beta);
            }
            if (!short_div) {
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    u[k] = div_step(u[k], nu[k], a.beta, a.rcp_beta);
                    v[k] = div_step(v[k], nv[k], a.beta, a.rcp_beta);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                u[k] = div_fixed(nu[k], a.beta, a.rcp_beta, fast);
                v[k] = div_fixed(nv[k], a.beta, a.rcp_beta, fast);
            }
        }
    }

    if (!LAST) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.mine(k, a.w, a.h)) a.ddst[st.cell(k, a.w)] = make_float2(u[k], v[k]);
        return;
    }

    // Divergence of D^K on the tile (src/fluid.cpp:221-237).  The plane
    // written here was last read in sweep NS-1, which every thread has left.
    float2 *B = plane + (NS & 1) * PL;
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (st.active) B[(st.r0 + k) * RX + st.x] = make_float2(u[k], v[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < S; ++k) {
        if (!st.mine(k, a.w, a.h)) continue;
        const int o = (st.r0 + k) * RX + st.x;
        const float uR = B[o + 1].x - B[o - 1].x;
        const float vT = B[o + RX].y - B[o - RX].y;
        const float b = a.gamma * (uR + vT);
        const int cell = st.cell(k, a.w);
        a.pw[cell] = warm[k];
        a.bpl[cell] = b;
        a.ddst[cell] = make_float2(a.k_even ? u[k] : pu[k], a.k_even ? v[k] : pv[k]);
    }
}

template <int H, bool LAST, int TS = kTile>
__device__ __forceinline__ void pressure_tile(const PressureArgs &a, int tile) {
    using G = Geo<H, TS>;
    constexpr int S = G::S, RY = G::RY;
    // The padded pitch is taken only where it keeps the workgroups that fit
    // a CU (160 KiB of LDS, 32 wavefronts): the planes plus the dye's
    // velocity block.
    constexpr int VN = LAST ? G::TX * G::TY * 8 : 8;
    constexpr int PP = wgs_per_cu(2 * G::RYP * G::P * 4 + VN, G::NT) >=
                               wgs_per_cu(2 * G::RYP * G::RX * 4 + VN, G::NT)
                           ? G::P
                           : G::RX;
    constexpr int PL = G::RYP * PP;
    constexpr int NS = H - (LAST ? 1 : 0);
    __shared__ float plane[2 * PL];

    const Strip<G> st(threadIdx.x, tile, a.tiles_x, a.w, a.h);
    float4 *__restrict__ P4 = reinterpret_cast<float4 *>(a.P);
    float4 *__restrict__ Q4 = reinterpret_cast<float4 *>(a.Q);

    // The iterate and right-hand side of the staged region, needed by the
    // first sweep.  E, needed only by the gradient, is loaded after the first
    // sweep, under the arithmetic of the rest.
    float p[S], rhs[S];
    {
        const float *__restrict__ src = a.first ? a.pw : a.psrc;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int g = st.row_ok(k) ? st.gj[k] * a.w + st.gi : 0;
            p[k] = src[g];
            rhs[k] = a.bpl[g];
        }
    }
    // Pressure Jacobi sweeps (src/fluid.cpp:240-257):
    // p' = (pL + pR + pT + pB + 1*b) / 4, left to right.
    // Sweeps are unrolled (no copies of the iterate between them) and the
    // last sweep of the last chunk is peeled, so that p^(K-1) is kept by
    // plain register copies rather than a select in every sweep.
    auto sweep = [&](int s) {
        float *B = plane + (s & 1) * PL;
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.active) B[(st.r0 + k) * PP + st.x] = p[k];
        __syncthreads();
        if (!st.active) return;
        const float top = B[(st.r0 > 0 ? st.r0 - 1 : 0) * PP + st.x];
        const float bot = B[(st.r0 + S < RY ? st.r0 + S : RY - 1) * PP + st.x];
        const float *Bl = B + st.r0 * PP + st.xl, *Br = B + st.r0 * PP + st.xr;
        float prev = top;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const float cur = p[k];
            const float below = k == S - 1 ? bot : p[k + 1];
            p[k] = pressure_cell(Bl[k * PP], Br[k * PP], prev, below, rhs[k]);
            prev = cur;
        }
    };
    float2 e[S];  // E of the tile cells (other rows load cell 0, unused)
    auto load_e = [&]() {
#pragma unroll
        for (int k = 0; k < S; ++k) e[k] = a.E[st.mine(k, a.w, a.h) ? st.cell(k, a.w) : 0];
    };
    float pp[S];  // p^(K-1)
    constexpr int NLOOP = LAST ? NS - 1 : NS;
#pragma unroll
    for (int s = 1; s <= NLOOP; ++s) {
        sweep(s);
        if (LAST && s == 1) load_e();
    }
    if (LAST) {
        if (NLOOP < 1) load_e();
#pragma unroll
        for (int k = 0; k < S; ++k) pp[k] = p[k];
        sweep(NS);
    }

    if (!LAST) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (st.mine(k, a.w, a.h)) a.pdst[st.cell(k, a.w)] = p[k];
        return;
    }

    // Gradient subtraction on the tile (src/fluid.cpp:273-295).
    float *B = plane + ((NS + 1) & 1) * PL;
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (st.active) B[(st.r0 + k) * PP + st.x] = p[k];
    __syncthreads();
    // The new velocities also go to LDS as they are made, for the dye pass.
    __shared__ float2 Vn[LAST ? G::TX * G::TY : 1];
    const bool dye = a.image != nullptr;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        if (!st.mine(k, a.w, a.h)) continue;
        const int o = (st.r0 + k) * PP + st.x;
        const float gradX = (B[o + 1] - B[o - 1]) * a.dt / 2.0f;
        const float gradY = (B[o + PP] - B[o - PP]) * a.dt / 2.0f;
        const int cell = st.cell(k, a.w);
        const float un = e[k].x - gradX, vn = e[k].y - gradY;
        P4[cell] = make_float4(un, vn, pp[k], rhs[k]);
        Q4[cell] = make_float4(e[k].x, e[k].y, p[k], rhs[k]);
        if (dye) Vn[(st.r0 + k - H) * G::TX + (st.x - H)] = make_float2(un, vn);
    }
    if (!dye) return;

    // Dye advection of the tile's pixels (src/fluid.cpp:89-124 with the
    // image the size of the field, so pixel (i, j) samples cell (i, j)),
    // pixels spread over all threads so that their gathers are in flight
    // together.  (Staging the dye tile in LDS during the sweeps, by the LDS
    // DMA path, measured slower: its burst competes with the sweeps.)
    __syncthreads();
    const float4 *__restrict__ I4 = reinterpret_cast<const float4 *>(a.image);
    float4 *__restrict__ O4 = reinterpret_cast<float4 *>(a.itmp);
    const int tx0 = st.x0 + H, ty0 = st.y0 + H;
    constexpr int PER = (G::TX * G::TY + G::NT - 1) / G::NT;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int idx = threadIdx.x + m * G::NT;
        if (idx >= G::TX * G::TY) break;
        const int gx = tx0 + idx % G::TX, gy = ty0 + idx / G::TX;
        if (gx >= a.w || gy >= a.h) continue;
        const float2 vel = Vn[idx];
        const float xp = (float)gx - div_n(a.dt_viw * vel.x, a.fw, a.inv_w, a.pow2_w);
        const float yp = (float)gy - div_n(a.dt_vih * vel.y, a.fh, a.inv_h, a.pow2_h);
        const Trace t = trace(xp, yp, a.w, a.h, a.fw, a.fh);
        const float4 f00 = I4[t.j0 * a.w + t.i0], f10 = I4[t.j0 * a.w + t.i1];
        const float4 f01 = I4[t.j1 * a.w + t.i0], f11 = I4[t.j1 * a.w + t.i1];
        O4[gy * a.w + gx] = make_float4(bilerp(t.sx, t.sy, f00.x, f10.x, f01.x, f11.x),
                                        bilerp(t.sx, t.sy, f00.y, f10.y, f01.y, f11.y),
                                        bilerp(t.sx, t.sy, f00.z, f10.z, f01.z, f11.z),
                                        bilerp(t.sx, t.sy, f00.w, f10.w, f01.w, f11.w));
    }
}

// ---- host-side helpers shared by the single-scene and batched launchers ----

// Workspace of one scene (floats, N = w*h): pw [0, N), b [N, 2N), the
// float2 planes d0 [2N, 4N), d1 [4N, 6N) and, for K > kMaxChunk, the
// pressure iterate planes p0 [6N, 7N), p1 [7N, 8N).  Chunk c of the
// diffusion writes d[c & 1]: the next chunk's input, or E when c is the
// last chunk (one chunk: d0 only).
inline size_t scene_ws_floats(int w, int h, int iters) {
    return (iters > kMaxChunk ? 8 : 4) * (size_t)w * (size_t)h;
}

// The launch sequence of one step: chunk c runs ns sweeps with halo H.
struct Chunk {
    int ns, H;
    bool first, last;
};
inline int chunk_plan(int iters, Chunk *out, int max_chunks, int chunk = kMaxChunk) {
    int n = 0;
    for (int done = 0; done < iters && n < max_chunks; ++n) {
        const int ns = (iters - done) < chunk ? (iters - done) : chunk;
        out[n].ns = ns;
        out[n].first = done == 0;
        out[n].last = done + ns == iters;
        out[n].H = ns + (out[n].last ? 1 : 0);
        done += ns;
    }
    return n;
}

// Scene-wide arguments of the two kernels (chunk fields set per launch).
inline DiffuseArgs diffuse_args(const float *P, const float *Q, int w, int h, float dt, float visc,
                                int iters, float *ws, int ts = kTile) {
    DiffuseArgs d{};
    const size_t N = (size_t)w * h;
    d.P = P;
    d.Q = Q;
    d.pw = ws;
    d.bpl = ws + N;
    d.w = w;
    d.h = h;
    d.tiles_x = (w + ts - 1) / ts;
    d.ntiles = d.tiles_x * ((h + ts - 1) / ts);
    d.dt = dt;
    d.fw = (float)w;
    d.fh = (float)h;
    d.inv_w = 1.0f / (float)w;
    d.inv_h = 1.0f / (float)h;
    d.pow2_w = (w & (w - 1)) == 0;
    d.pow2_h = (h & (h - 1)) == 0;
    // diffuse's coefficients, src/fluid.cpp:144-145 (beta evaluated in double)
    d.alpha = visc * dt;
    d.beta = (float)(1.0 + 4.0 * (double)d.alpha);
    d.rcp_beta = 1.0f / d.beta;
    d.fast_div = 0;
    if (d.beta >= 0x1p-20f && d.beta <= 0x1p20f) {  // false for NaN
        // y*beta is exact in double (24 x 24 bits).
        const double e = (double)d.rcp_beta * (double)d.beta - 1.0;
        d.fast_div = (e <= 0x1p-26 && e >= -0x1p-26) ? 2 : 1;
    }
    d.gamma = (float)(-1.0 / (double)dt);  // src/fluid.cpp:218
    d.k_even = (iters % 2) == 0;
    return d;
}

inline void diffuse_chunk(DiffuseArgs &d, int c, const Chunk &ch, float *ws) {
    const size_t N = (size_t)d.w * d.h;
    float2 *dpl[2] = {reinterpret_cast<float2 *>(ws + 2 * N),
                      reinterpret_cast<float2 *>(ws + 4 * N)};
    d.first = ch.first;
    d.dsrc = dpl[(c + 1) & 1];
    d.ddst = dpl[c & 1];
}

// `image`/`itmp` non-null: advect the dye (same size as the field, 4
// channels) in the last chunk.
inline PressureArgs pressure_args(float *P, float *Q, const float *image, float *itmp, int w,
                                  int h, float dt, float *ws, int ts = kTile) {
    PressureArgs p{};
    const size_t N = (size_t)w * h;
    p.P = P;
    p.Q = Q;
    p.pw = ws;
    p.bpl = ws + N;
    p.image = image;
    p.itmp = itmp;
    p.w = w;
    p.h = h;
    p.tiles_x = (w + ts - 1) / ts;
    p.ntiles = p.tiles_x * ((h + ts - 1) / ts);
    p.dt = dt;
    p.fw = (float)w;
    p.fh = (float)h;
    p.inv_w = 1.0f / (float)w;
    p.inv_h = 1.0f / (float)h;
    p.pow2_w = (w & (w - 1)) == 0;
    p.pow2_h = (h & (h - 1)) == 0;
    // src/fluid.cpp:82-83,97-98 with the image the size of the field.
    const float viw = (float)w / (float)w, vih = (float)h / (float)h;
    p.dt_viw = dt / viw;
    p.dt_vih = dt / vih;
    return p;
}

inline void pressure_chunk(PressureArgs &p, int c, const Chunk &ch, float *ws,
                           const float *image, float *itmp) {
    const size_t N = (size_t)p.w * p.h;
    float *ppl[2] = {ws + 6 * N, ws + 7 * N};
    p.E = reinterpret_cast<const float2 *>(ws + 2 * N + 2 * N * (c & 1));
    p.first = ch.first;
    p.psrc = ppl[(c + 1) & 1];
    p.pdst = ppl[c & 1];
    p.image = ch.last ? image : nullptr;
    p.itmp = ch.last ? itmp : nullptr;
}

}  // namespace pfs
