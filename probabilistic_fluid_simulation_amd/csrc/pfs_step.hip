// Fused, temporally blocked fluid step for gfx950.
//
// One call of the reference's simulate_fluid_step (src/fluid.cpp:298-310)
// is 2K+3 full-grid passes there.  Here it is two kernel groups:
//
//   D: advect (src/fluid.cpp:24-70) + K diffusion sweeps (:129-196) +
//      divergence (:221-237), temporally blocked in LDS;
//   P: K pressure sweeps (:239-266) + gradient subtraction (:269-296),
//      temporally blocked in LDS.
//
// Each workgroup owns a TX x TY tile of cells.  It stages the tile plus a
// halo of (sweeps + 1) cells in LDS (periodic wrap folded into the staged
// coordinates), runs the sweeps there on shrinking rectangles (ping-pong
// buffers, one barrier per sweep), and writes only its own tile.  Halo cells
// are recomputed by both neighbours with the same operations in the same
// order, so every value is bit-identical to the reference's global sweep.
// When K exceeds what fits in LDS the sweeps run in chunks of kMaxChunk,
// with intermediate iterates in workspace planes.
//
// Buffer contract (equal to fluid.cpp's, derived from its pointer swaps):
// with P = vp and Q = tmp at entry, E = D^K if K is even else D^(K-1),
// warm = Q[2] if K is even else P[2] (the pressure start), p^0 = warm:
//   Q = (E.u, E.v, p^K, b)     P = (E - grad p^K, p^(K-1), b)
// b = divergence of D^K.  The pressure start is copied to a workspace plane
// by group D, so group P's halo reads never race with its own writes.

#include "pfs_device.h"
#include "pfs_internal.h"

namespace pfs {

namespace {

constexpr int kTX = 64;
constexpr int kTY = 64;
constexpr int kNT = 1024;
constexpr int kMaxChunk = 8;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Remap so that each XCD owns a contiguous band of tiles and the
// halos its tiles share stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    if ((n & 7) != 0) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
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
    float alpha, beta, gamma;
    int ns, first, last, k_even;
};

struct PressureArgs {
    float *P;
    float *Q;
    const float *psrc;  // iterate plane read by a non-first chunk
    float *pdst;        // iterate plane written by a non-last chunk
    const float *pw;
    const float *bpl;
    int w, h, tiles_x, ntiles;
    float dt;
    int ns, first, last;
};

__device__ __forceinline__ int wrap_index(int g, int n) {
    g %= n;
    return g < 0 ? g + n : g;
}

template <int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_diffuse_chunk(DiffuseArgs a) {
    extern __shared__ float4 lds4[];
    const int H = a.ns + a.last;
    const int RX = TX + 2 * H, RY = TY + 2 * H, RN = RX * RY;
    float2 *buf0 = reinterpret_cast<float2 *>(lds4);
    float2 *buf1 = buf0 + RN;
    int *wx = reinterpret_cast<int *>(buf1 + RN);
    int *wy = wx + RX;

    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x0 = tx * TX - H, y0 = ty * TY - H;  // global coords of region (0, 0)
    for (int t = threadIdx.x; t < RX; t += NT) wx[t] = wrap_index(x0 + t, a.w);
    for (int t = threadIdx.x; t < RY; t += NT) wy[t] = wrap_index(y0 + t, a.h);
    __syncthreads();

    const float2 *P2 = reinterpret_cast<const float2 *>(a.P);  // (u, v) at index 2*cell
    if (a.first) {
        // Advected velocity (src/fluid.cpp:36-67) of every staged cell.
        for_rect<NT>(0, 0, RX, RY, [&](int rx, int ry) {
            const int gi = wx[rx], gj = wy[ry];
            const float2 uv = P2[2 * (gj * a.w + gi)];
            const float xp = (float)gi - div_n(a.dt * uv.x, a.fw, a.inv_w, a.pow2_w);
            const float yp = (float)gj - div_n(a.dt * uv.y, a.fh, a.inv_h, a.pow2_h);
            const Trace t = trace(xp, yp, a.w, a.h, a.fw, a.fh);
            const float2 f00 = P2[2 * (t.j0 * a.w + t.i0)];
            const float2 f10 = P2[2 * (t.j0 * a.w + t.i1)];
            const float2 f01 = P2[2 * (t.j1 * a.w + t.i0)];
            const float2 f11 = P2[2 * (t.j1 * a.w + t.i1)];
            buf0[ry * RX + rx] = make_float2(bilerp(t.sx, t.sy, f00.x, f10.x, f01.x, f11.x),
                                             bilerp(t.sx, t.sy, f00.y, f10.y, f01.y, f11.y));
        });
    } else {
        for_rect<NT>(0, 0, RX, RY, [&](int rx, int ry) {
            buf0[ry * RX + rx] = a.dsrc[wy[ry] * a.w + wx[rx]];
        });
    }
    __syncthreads();

    // Jacobi smoothing sweeps (src/fluid.cpp:154-185) on shrinking rectangles.
    for (int s = 1; s <= a.ns; ++s) {
        const float2 *in = (s & 1) ? buf0 : buf1;
        float2 *out = (s & 1) ? buf1 : buf0;
        for_rect<NT>(s, s, RX - 2 * s, RY - 2 * s, [&](int rx, int ry) {
            const int c = ry * RX + rx;
            const float2 C = in[c], L = in[c - 1], R = in[c + 1], T = in[c - RX],
                         B = in[c + RX];
            out[c] = make_float2(diffuse_cell(L.x, R.x, T.x, B.x, C.x, a.alpha, a.beta),
                                 diffuse_cell(L.y, R.y, T.y, B.y, C.y, a.alpha, a.beta));
        });
        __syncthreads();
    }
    const float2 *dk = (a.ns & 1) ? buf1 : buf0;    // after the chunk's last sweep
    const float2 *dkm1 = (a.ns & 1) ? buf0 : buf1;  // before it

    if (!a.last) {
        for_rect<NT>(H, H, TX, TY, [&](int rx, int ry) {
            const int gx = x0 + rx, gy = y0 + ry;
            if (gx < a.w && gy < a.h) a.ddst[gy * a.w + gx] = dk[ry * RX + rx];
        });
        return;
    }
    float4 *Q4 = reinterpret_cast<float4 *>(a.Q);
    for_rect<NT>(H, H, TX, TY, [&](int rx, int ry) {
        const int gx = x0 + rx, gy = y0 + ry;
        if (gx >= a.w || gy >= a.h) return;
        const int c = ry * RX + rx;
        const int cell = gy * a.w + gx;
        // Divergence of D^K (src/fluid.cpp:230-235).
        const float uR = dk[c + 1].x - dk[c - 1].x;
        const float vT = dk[c + RX].y - dk[c - RX].y;
        const float b = a.gamma * (uR + vT);
        const float2 E = a.k_even ? dk[c] : dkm1[c];
        const float4 q = Q4[cell];
        a.pw[cell] = a.k_even ? q.z : a.P[4 * cell + 2];
        a.bpl[cell] = b;
        Q4[cell] = make_float4(E.x, E.y, q.z, b);
    });
}

template <int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_pressure_chunk(PressureArgs a) {
    extern __shared__ float4 lds4[];
    const int H = a.ns + a.last;
    const int RX = TX + 2 * H, RY = TY + 2 * H, RN = RX * RY;
    float *buf0 = reinterpret_cast<float *>(lds4);
    float *buf1 = buf0 + RN;
    float *bb = buf1 + RN;
    int *wx = reinterpret_cast<int *>(bb + RN);
    int *wy = wx + RX;

    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x0 = tx * TX - H, y0 = ty * TY - H;
    for (int t = threadIdx.x; t < RX; t += NT) wx[t] = wrap_index(x0 + t, a.w);
    for (int t = threadIdx.x; t < RY; t += NT) wy[t] = wrap_index(y0 + t, a.h);
    __syncthreads();

    const float *src = a.first ? a.pw : a.psrc;
    for_rect<NT>(0, 0, RX, RY, [&](int rx, int ry) {
        const int g = wy[ry] * a.w + wx[rx];
        buf0[ry * RX + rx] = src[g];
        bb[ry * RX + rx] = a.bpl[g];
    });
    __syncthreads();

    // Pressure Jacobi sweeps (src/fluid.cpp:240-257).
    for (int s = 1; s <= a.ns; ++s) {
        const float *in = (s & 1) ? buf0 : buf1;
        float *out = (s & 1) ? buf1 : buf0;
        for_rect<NT>(s, s, RX - 2 * s, RY - 2 * s, [&](int rx, int ry) {
            const int c = ry * RX + rx;
            out[c] = pressure_cell(in[c - 1], in[c + 1], in[c - RX], in[c + RX], bb[c]);
        });
        __syncthreads();
    }
    const float *pk = (a.ns & 1) ? buf1 : buf0;
    const float *pkm1 = (a.ns & 1) ? buf0 : buf1;

    if (!a.last) {
        for_rect<NT>(H, H, TX, TY, [&](int rx, int ry) {
            const int gx = x0 + rx, gy = y0 + ry;
            if (gx < a.w && gy < a.h) a.pdst[gy * a.w + gx] = pk[ry * RX + rx];
        });
        return;
    }
    float4 *P4 = reinterpret_cast<float4 *>(a.P);
    float4 *Q4 = reinterpret_cast<float4 *>(a.Q);
    for_rect<NT>(H, H, TX, TY, [&](int rx, int ry) {
        const int gx = x0 + rx, gy = y0 + ry;
        if (gx >= a.w || gy >= a.h) return;
        const int c = ry * RX + rx;
        const int cell = gy * a.w + gx;
        // Gradient subtraction (src/fluid.cpp:282-293).
        const float gradX = (pk[c + 1] - pk[c - 1]) * a.dt / 2.0f;
        const float gradY = (pk[c + RX] - pk[c - RX]) * a.dt / 2.0f;
        const float4 q = Q4[cell];  // (E.u, E.v, start, b)
        const float b = bb[c];
        P4[cell] = make_float4(q.x - gradX, q.y - gradY, pkm1[c], b);
        Q4[cell] = make_float4(q.x, q.y, pk[c], b);
    });
}

size_t diffuse_lds(int ns, int last) {
    const int H = ns + last;
    const size_t RX = kTX + 2 * H, RY = kTY + 2 * H;
    return 2 * RX * RY * sizeof(float2) + (RX + RY) * sizeof(int);
}

size_t pressure_lds(int ns, int last) {
    const int H = ns + last;
    const size_t RX = kTX + 2 * H, RY = kTY + 2 * H;
    return 3 * RX * RY * sizeof(float) + (RX + RY) * sizeof(int);
}

int set_lds_limits() {
    static int done = 0;
    static int rc = PFS_OK;
    if (done) return rc;
    const size_t dmax = diffuse_lds(kMaxChunk, 1), pmax = pressure_lds(kMaxChunk, 1);
    hipError_t e1 = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_diffuse_chunk<kTX, kTY, kNT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)dmax);
    hipError_t e2 = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_pressure_chunk<kTX, kTY, kNT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)pmax);
    done = 1;
    if (e1 != hipSuccess || e2 != hipSuccess)
        rc = set_error(PFS_ERR_HIP, "hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
    return rc;
}

// Workspace layout (float offsets, N = w*h): pw [0, N), b [N, 2N),
// iterate planes d0 [2N, 4N), d1 [4N, 6N) (float2), p0 [6N, 7N), p1 [7N, 8N).
size_t ws_floats(int w, int h) { return 8 * (size_t)w * (size_t)h; }

}  // namespace

size_t fused_workspace_bytes(int w, int h) { return ws_floats(w, h) * sizeof(float); }

int fused_step(const StepArgs &s, float *ws, hipStream_t stream, hipEvent_t *ev) {
    PFS_TRY(set_lds_limits());
    const int w = s.w, h = s.h, K = s.iters;
    const size_t N = (size_t)w * h;
    const int tiles_x = (w + kTX - 1) / kTX, tiles_y = (h + kTY - 1) / kTY;
    const int ntiles = tiles_x * tiles_y;
    float *pw = ws, *bpl = ws + N;
    float2 *dpl[2] = {reinterpret_cast<float2 *>(ws + 2 * N), reinterpret_cast<float2 *>(ws + 4 * N)};
    float *ppl[2] = {ws + 6 * N, ws + 7 * N};
    const Pow2 p2w(w), p2h(h);
    const float alpha = diffuse_alpha(s.visc, s.dt);

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
    d.beta = diffuse_beta(alpha);
    d.gamma = pressure_gamma(s.dt);
    d.k_even = (K % 2) == 0;

    if (ev) PFS_HIP_TRY(hipEventRecord(ev[0], stream));
    int done = 0, chunk = 0;
    while (done < K) {
        const int ns = (K - done) < kMaxChunk ? (K - done) : kMaxChunk;
        d.ns = ns;
        d.first = (done == 0);
        d.last = (done + ns == K);
        d.dsrc = dpl[(chunk + 1) & 1];
        d.ddst = dpl[chunk & 1];
        hipLaunchKernelGGL((k_diffuse_chunk<kTX, kTY, kNT>), dim3(ntiles), dim3(kNT),
                           diffuse_lds(ns, d.last), stream, d);
        PFS_TRY(check_launch("fused diffuse chunk"));
        done += ns;
        ++chunk;
    }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[1], stream));

    PressureArgs p{};
    p.P = s.P;
    p.Q = s.Q;
    p.pw = pw;
    p.bpl = bpl;
    p.w = w;
    p.h = h;
    p.tiles_x = tiles_x;
    p.ntiles = ntiles;
    p.dt = s.dt;
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[2], stream));
    done = 0;
    chunk = 0;
    while (done < K) {
        const int ns = (K - done) < kMaxChunk ? (K - done) : kMaxChunk;
        p.ns = ns;
        p.first = (done == 0);
        p.last = (done + ns == K);
        p.psrc = ppl[(chunk + 1) & 1];
        p.pdst = ppl[chunk & 1];
        hipLaunchKernelGGL((k_pressure_chunk<kTX, kTY, kNT>), dim3(ntiles), dim3(kNT),
                           pressure_lds(ns, p.last), stream, p);
        PFS_TRY(check_launch("fused pressure chunk"));
        done += ns;
        ++chunk;
    }
    if (ev) PFS_HIP_TRY(hipEventRecord(ev[3], stream));
    (void)tiles_y;
    return PFS_OK;
}

}  // namespace pfs
