// Particle layer over the grid fluid (gfx950).  No counterpart in the
// reference, which has no particles: these are tracers carried by the grid's
// velocity field, with a uniform neighbour grid over them.
//
//   pfs_particles_advect     x += (dt*u(x))/W, y likewise (the reference's
//                            back-trace, src/fluid.cpp:39-41, run forward),
//                            u bilinear at the position with the reference's
//                            corner keys and blend order (:55-66); positions
//                            wrap back into [0, W) exactly;
//   pfs_ngrid_build          cell keys; a stable LSD radix sort of
//                            (key, particle) pairs, 8- or 10-bit digits: per-block
//                            LDS histograms, a DPP prefix scan, and a scatter
//                            whose in-block ranks come from wavefront ballots;
//                            cell start/end tables (the scan of the cell
//                            counts: every cell, empty or not, gets its
//                            lower bound); positions in sorted order;
//   pfs_ngrid_neighbors      neighbour counts, their exclusive scan, and the
//                            lists (CSR) of every particle within radius r,
//                            visiting the 3x3 cells in a fixed order;
//   pfs_ngrid_density        SPH density from the lists' traversal with the
//                            poly6 kernel, FP32 sums in traversal order.
//
// Every integer output (keys, sort order, tables, lists) is deterministic
// and equals the CPU restatement in oracle/pfs_particles.c bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "pfs.h"
#include "pfs_device.h"
#include "pfs_internal.h"

namespace pfs {
namespace {

constexpr int kThreads = 1024;              // sort, scan and grid kernels
constexpr int kItems = 4;                   // keys per thread
constexpr int kTile = kThreads * kItems;    // keys per block
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRadixBits = 10;           // digits of 8 or 10 bits

// ---- wavefront and block scans ------------------------------------------
// Inclusive scan across the 64 lanes: four row shifts (1, 2, 4, 8) give
// every 16-lane row its prefix, then the row totals are carried by the
// row_bcast:15 (into rows 1, 3) and row_bcast:31 (into rows 2, 3) DPP forms.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Exclusive scan over the block's kThreads threads; *total gets the sum.
// Uses `sh` (kWaves + 1 words of LDS).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *sh, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(x);
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = lane < kWaves ? sh[lane] : 0u;
        const uint32_t wi = wave_incl_scan(w);
        if (lane < kWaves) sh[lane] = wi - w;  // exclusive over waves
        if (lane == kWaves - 1) sh[kWaves] = wi;
    }
    __syncthreads();
    const uint32_t out = sh[wave] + inc - x;
    *total = sh[kWaves];
    __syncthreads();  // sh may be reused by the caller
    return out;
}

// ---- generic exclusive scan (three kernels) --------------------------------
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t *__restrict__ in, int n,
                                                          uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[kWaves + 1];
    const size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < (size_t)n) s += in[base + k];
    uint32_t total;
    (void)block_excl_scan(s, sh, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Block b adds up the totals of blocks [0, b) itself (a few hundred at
// most), then scans its tile; the last block also writes the grand total.
__global__ __launch_bounds__(kThreads) void k_scan_down(const uint32_t *__restrict__ in,
                                                        uint32_t *__restrict__ out, int n,
                                                        const uint32_t *__restrict__ sums,
                                                        uint32_t *__restrict__ grand) {
    __shared__ uint32_t sh[kWaves + 1];
    uint32_t before = 0;
    for (int j = threadIdx.x; j < (int)blockIdx.x; j += kThreads) before += sums[j];
    uint32_t offset;
    (void)block_excl_scan(before, sh, &offset);
    const size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems], s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = base + k < (size_t)n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = offset + block_excl_scan(s, sh, &total);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (base + k < (size_t)n) out[base + k] = run;
        run += v[k];
    }
    if (grand && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *grand = offset + total;
}

int nblocks_for(size_t n) { return (int)((n + kTile - 1) / kTile); }

// Exclusive scan of n values in two kernels (block totals, then blocks
// scanned from the totals before them); grand (device, may be null) gets the
// total.  ws: nblocks_for(n) words.
int scan_exclusive(const uint32_t *in, uint32_t *out, int n, uint32_t *grand, uint32_t *ws,
                   hipStream_t s) {
    const int nb = nblocks_for((size_t)n);
    if (nb == 0) return PFS_OK;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kThreads), 0, s, in, n, ws);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kThreads), 0, s, in, out, n, ws, grand);
    return check_launch("exclusive scan");
}

// ---- stable LSD radix sort of (key, value) pairs ---------------------------
// Digits of RB bits (8 or 10).  The histogram of block b is stored digit
// major, hist[d * nb + b]; each digit's row is then scanned on its own
// (k_digit_rows), and the scatter adds the digit's base, the exclusive scan
// of the row totals, which every scatter block forms for itself.

// Cell of a position in [0, W): (int)(x * kx) with kx = gx / W (float),
// clamped into the grid.
__device__ __forceinline__ int cell_of(float x, float k, int g) {
    const int c = (int)(x * k);
    return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

struct CellKeyArgs {  // the first pass computes the keys from the positions
    const float *x, *y;
    float kx, ky;
    int gx, gy;
};

// Block histograms of one digit.  With FROM_XY the keys are made here from
// the positions, cy * gx + cx, and written to `keys`.
template <int RB, bool FROM_XY>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint32_t *__restrict__ kin,
                                                         CellKeyArgs ck,
                                                         uint32_t *__restrict__ keys, int n,
                                                         int shift, uint32_t *__restrict__ hist,
                                                         int nb) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t h[R];
    for (int d = threadIdx.x; d < R; d += kThreads) h[d] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const size_t i = base + (size_t)r * kThreads + threadIdx.x;
        if (i < (size_t)n) {
            uint32_t k;
            if (FROM_XY) {
                k = (uint32_t)(cell_of(ck.y[i], ck.ky, ck.gy) * ck.gx + cell_of(ck.x[i], ck.kx, ck.gx));
                keys[i] = k;
            } else {
                k = kin[i];
            }
            atomicAdd(&h[(k >> shift) & (R - 1)], 1u);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < R; d += kThreads) hist[(size_t)d * nb + blockIdx.x] = h[d];
}

// Block d scans row d of the digit-major histogram (nb block counts)
// exclusively into the same row of `rows`, and writes the row's total to
// tot[d].
__global__ __launch_bounds__(kThreads) void k_digit_rows(const uint32_t *__restrict__ hist,
                                                         uint32_t *__restrict__ rows,
                                                         uint32_t *__restrict__ tot, int nb) {
    __shared__ uint32_t sh[kWaves + 1];
    const size_t row = (size_t)blockIdx.x * nb;
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kTile) {
        const int j = b0 + threadIdx.x * kItems;
        uint32_t v[kItems], s = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            v[k] = j + k < nb ? hist[row + j + k] : 0u;
            s += v[k];
        }
        uint32_t total;
        uint32_t run = carry + block_excl_scan(s, sh, &total);
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (j + k < nb) rows[row + j + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup i to XCD i % 8);
// this numbers tiles so that each XCD takes a contiguous run of them.  Tiles
// next to each other write the two ends of the same output cache lines, which
// then meet in one L2.  A bijection on [0, nb) for any nb.
__device__ __forceinline__ int xcd_block(int i, int nb) {
    const int q = nb >> 3, r = nb & 7, x = i & 7;
    return x * q + min(x, r) + (i >> 3);
}

// Scatter of one pass.  Wave w of block b takes the kItems * 64 keys
// [b * kTile + w * 256, + 256) in kItems rounds of 64, so the tile's order is
// (wave, round, lane).  The rank of an item among equal digits is: the
// digit's base (scan of the row totals) plus its row scan at this block, plus
// the equal digits of earlier waves (per-wave LDS counts, scanned across
// waves once), plus those of this wave's earlier rounds (the wave's running
// count) and of the lanes below it (RB ballots, one per digit bit,
// intersected).  vin null: the values are the key indices.  xs non-null (the
// last pass): the positions are gathered into sorted order as well.
template <int RB>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout,
    uint32_t *__restrict__ vout, int n, int shift, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ tot, int nb, const float *__restrict__ x,
    const float *__restrict__ y, float *__restrict__ xs, float *__restrict__ ys) {
    constexpr int R = 1 << RB;
    static_assert(R <= kThreads, "one digit per thread");
    __shared__ uint32_t wcount[kWaves][R];
    __shared__ uint32_t gofs[R], loff[R];
    __shared__ uint32_t lkey[kTile], lval[kTile];
    __shared__ uint32_t sh[kWaves + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tb = xcd_block(blockIdx.x, nb);  // the tile of this workgroup
    for (int e = threadIdx.x; e < kWaves * R; e += kThreads) (&wcount[0][0])[e] = 0;
    {
        const uint32_t t = threadIdx.x < R ? tot[threadIdx.x] : 0u;
        uint32_t all;
        const uint32_t dbase = block_excl_scan(t, sh, &all);  // its barriers also order the clear
        if (threadIdx.x < R) gofs[threadIdx.x] = dbase + rows[(size_t)threadIdx.x * nb + tb];
    }
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t base = (size_t)tb * kTile + (size_t)wave * (kItems * 64) + lane;
    uint32_t key[kItems], val[kItems], dig[kItems], local[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const size_t i = base + (size_t)r * 64;
        const bool valid = i < (size_t)n;
        key[r] = valid ? kin[i] : 0u;
        val[r] = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const bool valid = base + (size_t)r * 64 < (size_t)n;
        const uint32_t d = (key[r] >> shift) & (R - 1);
        dig[r] = d;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        if (!valid) peers = 0;
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        // The wave's own running count of digit d: read by every lane of
        // the digit, advanced by its lowest lane.  A wave's LDS accesses
        // complete in order, and the wave barrier keeps the compiler from
        // moving the next round's read above this write.
        const uint32_t before = valid ? wcount[wave][d] : 0u;
        local[r] = before + rank;
        if (valid && rank == 0) wcount[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    uint32_t dtotal = 0;
    if (threadIdx.x < R) {  // digit counts of the waves -> exclusive across waves
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wcount[w][threadIdx.x];
            wcount[w][threadIdx.x] = dtotal;
            dtotal += c;
        }
    }
    // The tile is first put in digit order in LDS (loff[d]: the tile's first
    // slot of digit d), then written out by consecutive threads, so that each
    // store instruction covers a few runs of consecutive positions.
    {
        uint32_t all;
        const uint32_t lo = block_excl_scan(dtotal, sh, &all);
        if (threadIdx.x < R) loff[threadIdx.x] = lo;
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        if (base + (size_t)r * 64 >= (size_t)n) continue;
        const uint32_t li = loff[dig[r]] + wcount[wave][dig[r]] + local[r];
        if (li < (uint32_t)kTile) {  // always: kept as a bound
            lkey[li] = key[r];
            lval[li] = val[r];
        }
    }
    __syncthreads();
    const int count = (int)min((size_t)kTile, (size_t)n - (size_t)tb * kTile);
    for (int j = threadIdx.x; j < count; j += kThreads) {
        const uint32_t k = lkey[j], v = lval[j];
        const uint32_t d = (k >> shift) & (R - 1);
        const uint32_t pos = gofs[d] + ((uint32_t)j - loff[d]);
        if (pos < (uint32_t)n) {  // always, when the scans are right: kept as a bound
            kout[pos] = k;
            vout[pos] = v;
            if (xs && v < (uint32_t)n) {
                xs[pos] = x[v];
                ys[pos] = y[v];
            }
        }
    }
}

// ---- grid ------------------------------------------------------------------
// start[c] = the lower bound of cell c in the sorted keys (binary search;
// the top levels of every search hit the same few cache lines), which is
// also end[c - 1]; the last cell ends at n.
__global__ void k_cell_tables(const uint32_t *__restrict__ sk, int n, uint32_t *__restrict__ start,
                              uint32_t *__restrict__ end, uint32_t ncells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    uint32_t lo = 0, hi = (uint32_t)n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sk[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    start[c] = lo;
    if (c > 0) end[c - 1] = lo;
    if (c + 1 == ncells) end[c] = (uint32_t)n;
}

__global__ void k_gather_u32(const uint32_t *__restrict__ index, const uint32_t *__restrict__ in,
                             uint32_t *__restrict__ out, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t i = index[p];
    out[p] = i < (uint32_t)n ? in[i] : 0u;
}

// x_q - x_p on a periodic axis of length L (half = L/2), positions in
// [0, L): the nearer image.  Across the wrap the far operand is moved by L
// first, exactly (it lies in (L/2, L), so x - L is exact by Sterbenz), and
// the two nearby values are then subtracted with one rounding at their own
// scale; subtracting first would round at L's scale.
__device__ __forceinline__ float periodic_diff(float xq, float xp, float L, float half) {
    const float d = xq - xp;
    const float over = (xq - L) - xp, under = xq - (xp - L);  // both formed: selects, no branches
    return d > half ? over : (d < -half ? under : d);
}

struct NArgs {
    const float *xs, *ys;
    const uint32_t *order;  // null: outputs indexed by sorted position
    const uint32_t *sk, *start, *end;
    int n, gx, gy;
    float W, H, half_w, half_h, r2;
    __device__ uint32_t id(uint32_t p) const { return order ? order[p] : p; }
};

// Visits every particle q of the 3x3 cells around p's cell, p included (rows
// then columns, -1..1, then q ascending): f(q, d2) with d2 the squared
// periodic distance, for every candidate, so that callers accumulate without
// branching.  The cells of one row are consecutive keys, so the three are
// one range of sorted positions, [start of the left cell, end of the right
// cell), split in two where the row wraps.
template <typename F>
__device__ __forceinline__ void for_neighbors(const NArgs &a, int p, F &&f) {
    const uint32_t c = a.sk[p];
    if (c >= (uint32_t)a.gx * (uint32_t)a.gy) return;  // a cell by construction: kept as a bound
    const int cx = (int)(c % (uint32_t)a.gx), cy = (int)(c / (uint32_t)a.gx);
    const float px = a.xs[p], py = a.ys[p];
    // Candidates are taken four at a time, their loads issued together, and
    // visited in ascending q as one at a time would.  Away from the domain's
    // edges no candidate is more than 3 cells away, less than half the
    // domain when it spans 7 or more cells: there the periodic correction is
    // the identity and that loop skips it.
    auto scan = [&](uint32_t q0, uint32_t q1, auto &&d2_of) {
        q1 = min(q1, (uint32_t)a.n);
        uint32_t q = q0;
        const float *xp = a.xs + q0, *yp = a.ys + q0;
        for (; q + 4 <= q1; q += 4, xp += 4, yp += 4) {
            const float x0 = xp[0], x1 = xp[1], x2 = xp[2], x3 = xp[3];
            const float y0 = yp[0], y1 = yp[1], y2 = yp[2], y3 = yp[3];
            const float e0 = d2_of(x0, y0), e1 = d2_of(x1, y1), e2 = d2_of(x2, y2),
                        e3 = d2_of(x3, y3);
            f(q, e0);
            f(q + 1, e1);
            f(q + 2, e2);
            f(q + 3, e3);
        }
        for (; q < q1; ++q, ++xp, ++yp) f(q, d2_of(*xp, *yp));
    };
    auto rows = [&](auto &&d2_of) {
        for (int oy = -1; oy <= 1; ++oy) {
            const int ny = cy + oy < 0 ? cy + oy + a.gy : (cy + oy >= a.gy ? cy + oy - a.gy : cy + oy);
            const uint32_t row = (uint32_t)(ny * a.gx);
            if (cx == 0) {  // cells gx-1 | 0, 1
                scan(a.start[row + a.gx - 1], a.end[row + a.gx - 1], d2_of);
                scan(a.start[row], a.end[row + 1], d2_of);
            } else if (cx == a.gx - 1) {  // cells gx-2, gx-1 | 0
                scan(a.start[row + cx - 1], a.end[row + cx], d2_of);
                scan(a.start[row], a.end[row], d2_of);
            } else {
                scan(a.start[row + cx - 1], a.end[row + cx + 1], d2_of);
            }
        }
    };
    const bool big = a.gx >= 7 && a.gy >= 7;
    const bool interior = cx >= 1 && cx <= a.gx - 2 && cy >= 1 && cy <= a.gy - 2 && big;
    if (interior) {
        rows([&](float ox, float oy) {
            const float dx = ox - px, dy = oy - py;
            return dx * dx + dy * dy;
        });
    } else if (big) {
        // Near an edge of a grid of 7 or more cells a side, whether a
        // candidate lies across the wrap is known per range: cell gx-1
        // seen from cell 0 (its candidates lie in (W/2, W): shifted by -W),
        // cell 0 seen from cell gx-1 (p lies there: p shifted by -W), and
        // the same for rows.  Every other candidate is within 2W/7 of p and
        // every shifted one at least 5W/7 away before the shift, so these
        // are exactly the cases periodic_diff corrects, with its values:
        // (x_q + -W) - x_p and (x_q + -0) - (x_p - W).
        for (int oy = -1; oy <= 1; ++oy) {
            const int ny = cy + oy < 0 ? cy + oy + a.gy : (cy + oy >= a.gy ? cy + oy - a.gy : cy + oy);
            const uint32_t row = (uint32_t)(ny * a.gx);
            const float qsy = (cy == 0 && oy == -1) ? -a.H : -0.0f;
            const float pyu = (cy == a.gy - 1 && oy == 1) ? py - a.H : py;
            auto shifted = [&](float qsx, float pxu) {
                return [=](float ox, float oy_) {
                    const float dx = (ox + qsx) - pxu, dy = (oy_ + qsy) - pyu;
                    return dx * dx + dy * dy;
                };
            };
            if (cx == 0) {  // cells gx-1 | 0, 1
                scan(a.start[row + a.gx - 1], a.end[row + a.gx - 1], shifted(-a.W, px));
                scan(a.start[row], a.end[row + 1], shifted(-0.0f, px));
            } else if (cx == a.gx - 1) {  // cells gx-2, gx-1 | 0
                scan(a.start[row + cx - 1], a.end[row + cx], shifted(-0.0f, px));
                scan(a.start[row], a.end[row], shifted(-0.0f, px - a.W));
            } else {
                scan(a.start[row + cx - 1], a.end[row + cx + 1], shifted(-0.0f, px));
            }
        }
    } else {
        rows([&](float ox, float oy) {
            const float dx = periodic_diff(ox, px, a.W, a.half_w);
            const float dy = periodic_diff(oy, py, a.H, a.half_h);
            return dx * dx + dy * dy;
        });
    }
}

__global__ void k_nbr_count(NArgs a, uint32_t *__restrict__ counts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    uint32_t c = 0;
    for_neighbors(a, p, [&](uint32_t q, float d2) { c += d2 < a.r2 && (int)q != p ? 1u : 0u; });
    const uint32_t i = a.id((uint32_t)p);
    if (i < (uint32_t)a.n) counts[i] = c;
}

__global__ void k_nbr_fill(NArgs a, const uint32_t *__restrict__ offsets, uint32_t *__restrict__ nbr,
                           size_t capacity) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const uint32_t i = a.id((uint32_t)p);
    if (i >= (uint32_t)a.n) return;
    size_t o = offsets[i];
    for_neighbors(a, p, [&](uint32_t q, float d2) {
        if (d2 < a.r2 && (int)q != p) {
            if (o < capacity) nbr[o] = a.id(q);
            ++o;
        }
    });
}

// The same lists by sorted position (order null), written through LDS: the
// lists of a wave's 64 consecutive particles are one contiguous block of the
// output, so each wave gathers its entries in an LDS window of kFillWin
// slots and then writes the window out with consecutive lanes (a lane's
// own list would otherwise go out as one scattered word per store).  A
// block larger than the window takes further traversals, one per window.
// Waves work independently: the LDS handoff is within one wave, whose LDS
// accesses complete in order.  3072 slots hold a wave's lists at about 48
// neighbours per particle in one traversal (48 KB per workgroup).
constexpr int kFillWin = 3072;
__global__ __launch_bounds__(256) void k_nbr_fill_lds(NArgs a, const uint32_t *__restrict__ offsets,
                                                      const uint32_t *__restrict__ counts,
                                                      uint32_t *__restrict__ nbr, size_t capacity) {
    __shared__ uint32_t win[4][kFillWin];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int pw = blockIdx.x * blockDim.x + wave * 64;  // the wave's first particle
    if (pw >= a.n) return;                              // the whole wave is past the end
    const int plast = min(pw + 63, a.n - 1);
    const size_t base = offsets[pw], end = (size_t)offsets[plast] + counts[plast];
    uint32_t *buf = win[wave];
    for (size_t w0 = base; w0 < end; w0 += kFillWin) {
        if (p < a.n) {
            size_t o = offsets[p];
            for_neighbors(a, p, [&](uint32_t q, float d2) {
                if (d2 < a.r2 && (int)q != p) {
                    if (o >= w0 && o < w0 + kFillWin) buf[o - w0] = q;
                    ++o;
                }
            });
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const size_t stop = end - w0 < (size_t)kFillWin ? end - w0 : (size_t)kFillWin;
        for (size_t j = lane; j < stop; j += 64)
            if (w0 + j < capacity) nbr[w0 + j] = buf[j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// SPH density with the 2D poly6 kernel, W(d) = c (r^2 - d^2)^3; with counts,
// the neighbour counts of the same traversal.
// rho = mc * S with mc = mass * 4/(pi r^8) and S the FP32 sum, in traversal
// order and p included where it falls, of (r^2 - d2)^3 over the candidates
// with d2 < r^2.  A rejected candidate adds +0 (t clamped at 0), which leaves
// S unchanged, so the loop does not branch.  All terms are positive: S is
// within (m - 1) u of the exact sum of its m rounded terms.
__global__ void k_density(NArgs a, float mc, float *__restrict__ rho,
                          uint32_t *__restrict__ counts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    float sum = 0.0f;
    uint32_t cnt = 0;
    for_neighbors(a, p, [&](uint32_t, float d2) {
        const float t = fmaxf(a.r2 - d2, 0.0f);
        sum += t * t * t;
        cnt += d2 < a.r2 ? 1u : 0u;
    });
    const uint32_t i = a.id((uint32_t)p);
    if (i < (uint32_t)a.n) {
        rho[i] = mc * sum;
        // A finite p met itself (d2 = 0) and counted it; a non-finite p
        // counted nothing (every d2 it forms is NaN or Inf), as k_nbr_count.
        if (counts) counts[i] = cnt ? cnt - 1u : 0u;
    }
}

// A position wrapped into [0, n): unchanged when already inside (the
// reference's fmod(fmod(x,n)+n,n) would re-round it), else by the exact
// fmod; a tiny negative x whose x + n rounds to n becomes 0.
__device__ __forceinline__ float wrap_pos(float x, float n) {
    if (x >= 0.0f && x < n) return x;
    x = fmodf(x, n);
    if (x < 0.0f) x += n;
    return x >= n ? 0.0f : x;
}

// Tracers: one forward step through the velocity field, sampled bilinearly
// at the position: corners i0 = (int)x, i1 = (i0 + 1) mod W, weights and
// blend as src/fluid.cpp:55-66.
__global__ void k_tracer_advect(float *__restrict__ x, float *__restrict__ y, int n,
                                const float *__restrict__ vp, int w, int h, int vc, float dt,
                                float fw, float fh, float inv_w, float inv_h, int pow2_w,
                                int pow2_h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float px = x[i], py = y[i];
    const int i0 = min(max((int)px, 0), w - 1), j0 = min(max((int)py, 0), h - 1);
    const int i1 = i0 + 1 == w ? 0 : i0 + 1, j1 = j0 + 1 == h ? 0 : j0 + 1;
    const float sx = px - (float)i0, sy = py - (float)j0;
    const float *f00 = vp + ((size_t)j0 * w + i0) * vc, *f10 = vp + ((size_t)j0 * w + i1) * vc;
    const float *f01 = vp + ((size_t)j1 * w + i0) * vc, *f11 = vp + ((size_t)j1 * w + i1) * vc;
    const float u = bilerp(sx, sy, f00[0], f10[0], f01[0], f11[0]);
    const float v = bilerp(sx, sy, f00[1], f10[1], f01[1], f11[1]);
    x[i] = wrap_pos(px + div_n(dt * u, fw, inv_w, pow2_w), fw);
    y[i] = wrap_pos(py + div_n(dt * v, fh, inv_h, pow2_h), fh);
}

// ---- position-based fluids: the density constraint, Jacobi iterations ------
// Macklin & Mueller, "Position Based Fluids" (ACM TOG 32(4), 2013), in 2D,
// on the sorted positions, with the neighbour lists found at the start of
// the solve (as the paper does once per step), kernels evaluated at the
// current iterate:
//   C_i = rho_i / rho0 - 1,  rho_i = m * W_poly6 summed over i and its list
//   lambda_i = -C_i / (sum_k |grad_k C_i|^2 + eps)
//   dx_i = (1 / rho0) * sum_{j in list} (lambda_i + lambda_j) grad W_spiky(x_i - x_j)
// with W_poly6 = 4/(pi h^8) (h^2 - r^2)^3 and grad W_spiky(d) =
// -30/(pi h^5) (h - r)^2 d / r for 0 < r < h (zero outside).  All updates
// of an iteration read the same iterate (Jacobi); positions wrap.
struct PbfCtl {
    int done, final_it;  // converged flag; the iterate the solve ends on
};

struct PbfArgs {
    const float *xs, *ys;  // the current iterate (sorted positions)
    const uint32_t *noff, *ncnt, *nbr;  // neighbour lists by sorted position
    const uint32_t *sk, *cstart, *cend;  // the start's sorted keys and cell tables
    int gx, gy;
    int n;
    float W, H, half_w, half_h;
    float h, h2, mc, gc, rho0, r0sq, eps;
    float *lambda;   // [n] by sorted position
    float *nx, *ny;  // the next iterate
    double *part;    // [max_iters + 1][nblocks]: per-block sums of C^2
    double *res;     // [max_iters + 1]: sum of C^2 of each iterate
    PbfCtl *ctl;
    int it, nblocks;
    double tol2n;    // convergence: sum C^2 <= tol^2 * n
};

// The positions a workgroup's 256 particles gather, staged in LDS.  Their
// lists lie in the three row windows of their cells' 3x3 blocks: when the
// particles share one cell row away from the domain's edges, row cy + oy
// contributes the sorted positions [start(cy + oy, cx_first - 1),
// end(cy + oy, cx_last + 1)), a few hundred in all.  Staged is uniform over
// the workgroup; otherwise (several rows, an edge, or more than kStage
// positions) the gathers go to memory.
constexpr int kStage = 2048;
struct Staged {
    float x[kStage], y[kStage], l[kStage];
    uint32_t w0, w1, w2, len0, len01, len012;  // window starts; cumulative lengths
    int on;
};

template <bool WITH_LAMBDA>
__device__ __forceinline__ void stage_windows(const PbfArgs &b, Staged &st) {
    if (threadIdx.x == 0) {
        const int p0 = blockIdx.x * blockDim.x;
        const int pl = min(p0 + (int)blockDim.x, b.n) - 1;
        const uint32_t c0 = b.sk[p0], c1 = b.sk[pl];
        const int cy = (int)(c0 / (uint32_t)b.gx), cx0 = (int)(c0 % (uint32_t)b.gx),
                  cx1 = (int)(c1 % (uint32_t)b.gx);
        int on = (int)(c1 / (uint32_t)b.gx) == cy && cy >= 1 && cy <= b.gy - 2 && cx0 >= 1 &&
                 cx1 <= b.gx - 2 && c0 < (uint32_t)b.gx * (uint32_t)b.gy &&
                 c1 < (uint32_t)b.gx * (uint32_t)b.gy;
        uint32_t ws[3] = {0, 0, 0}, len[3] = {0, 0, 0};
        if (on) {
            for (int k = 0; k < 3; ++k) {
                const uint32_t row = (uint32_t)((cy - 1 + k) * b.gx);
                ws[k] = b.cstart[row + cx0 - 1];
                const uint32_t we = b.cend[row + cx1 + 1];
                len[k] = we > ws[k] ? we - ws[k] : 0u;
            }
            on = len[0] + len[1] + len[2] <= (uint32_t)kStage;
        }
        st.w0 = ws[0];
        st.w1 = ws[1];
        st.w2 = ws[2];
        st.len0 = len[0];
        st.len01 = len[0] + len[1];
        st.len012 = len[0] + len[1] + len[2];
        st.on = on;
    }
    __syncthreads();
    if (!st.on) return;
    // Window k's positions go to slots [len_0 + .. + len_(k-1), + len_k).
    for (uint32_t j = threadIdx.x; j < st.len012; j += blockDim.x) {
        uint32_t q;
        if (j < st.len0) q = st.w0 + j;
        else if (j < st.len01) q = st.w1 + (j - st.len0);
        else q = st.w2 + (j - st.len01);
        if (q < (uint32_t)b.n) {
            st.x[j] = b.xs[q];
            st.y[j] = b.ys[q];
            if (WITH_LAMBDA) st.l[j] = b.lambda[q];
        }
    }
    __syncthreads();
}

// The LDS slot of sorted position q, or kStage when q is in no window (the
// lists then did not come from these cells; such entries read memory).
__device__ __forceinline__ uint32_t staged_slot(const Staged &st, uint32_t q) {
    if (q >= st.w0 && q - st.w0 < st.len0) return q - st.w0;
    if (q >= st.w1 && q - st.w1 < st.len01 - st.len0) return q - st.w1 + st.len0;
    if (q >= st.w2 && q - st.w2 < st.len012 - st.len01) return q - st.w2 + st.len01;
    return (uint32_t)kStage;
}

// Every listed neighbour q of p, in list order: f(q, dx, dy, d2, slot) with
// (dx, dy) = x_q - x_p, periodic (the lists were found at the start; the
// positions have moved since) and slot q's LDS slot (kStage: not staged).
// Loads are issued eight at a time.
template <typename F>
__device__ __forceinline__ void for_listed(const PbfArgs &b, const Staged &st, int p, F &&f) {
    const float px = b.xs[p], py = b.ys[p];
    const uint32_t o = b.noff[p], c = b.ncnt[p];
    const bool on = st.on;
    auto one = [&](uint32_t q, uint32_t slot, float ox, float oy) {
        const float dx = periodic_diff(ox, px, b.W, b.half_w);
        const float dy = periodic_diff(oy, py, b.H, b.half_h);
        f(q, slot, dx, dy, dx * dx + dy * dy);
    };
    constexpr int kB = 8;  // neighbours whose loads are in flight together
    uint32_t k = 0;
    for (; k + kB <= c; k += kB) {
        uint32_t q[kB], sl[kB];
        float x[kB], y[kB];
#pragma unroll
        for (int j = 0; j < kB; ++j) q[j] = b.nbr[o + k + j];
#pragma unroll
        for (int j = 0; j < kB; ++j) {
            sl[j] = on ? staged_slot(st, q[j]) : (uint32_t)kStage;
            const bool lds = sl[j] < (uint32_t)kStage;
            x[j] = lds ? st.x[lds ? sl[j] : 0] : b.xs[q[j]];
            y[j] = lds ? st.y[lds ? sl[j] : 0] : b.ys[q[j]];
        }
#pragma unroll
        for (int j = 0; j < kB; ++j) one(q[j], sl[j], x[j], y[j]);
    }
    for (; k < c; ++k) {
        const uint32_t q = b.nbr[o + k];
        const uint32_t sl = on ? staged_slot(st, q) : (uint32_t)kStage;
        const bool lds = sl < (uint32_t)kStage;
        one(q, sl, lds ? st.x[lds ? sl : 0] : b.xs[q], lds ? st.y[lds ? sl : 0] : b.ys[q]);
    }
}

// w with grad W_spiky(x_p - x_q) = -w * (dx, dy): ((gc * t) * t) / r,
// t = h - r, r = sqrt(d2), for 0 < d2 < h^2; else 0 (selected, so that the
// loops do not branch; adding +-0 leaves the sums unchanged).
__device__ __forceinline__ float spiky_w(const PbfArgs &b, float d2) {
    const float r = sqrtf(d2);
    const float t = b.h - r;
    const float w = ((b.gc * t) * t) / r;
    return (d2 > 0.0f && d2 < b.h2) ? w : 0.0f;
}

// Block-wide FP64 sum (thread 0 gets it): wave shuffles, then the waves in order.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += part[k];
    return s;
}

// lambda of iterate b.it and the block sums of C^2 (256 threads).
__global__ __launch_bounds__(256) void k_pbf_lambda(PbfArgs b) {
    if (b.ctl->done) return;  // the same verdict for every thread
    __shared__ Staged st;
    stage_windows<false>(b, st);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double c2 = 0.0;
    if (p < b.n) {
        float S = b.h2 * b.h2 * b.h2;  // p itself
        float gx = 0.0f, gy = 0.0f, g2 = 0.0f;
        for_listed(b, st, p, [&](uint32_t, uint32_t, float dx, float dy, float d2) {
            const float t2 = fmaxf(b.h2 - d2, 0.0f);
            S += t2 * t2 * t2;
            const float w = spiky_w(b, d2);
            const float vx = w * dx, vy = w * dy;
            gx += vx;
            gy += vy;
            g2 += vx * vx + vy * vy;
        });
        const float rho = b.mc * S;
        const float C = rho / b.rho0 - 1.0f;
        const float denom = (gx * gx + gy * gy + g2) / b.r0sq + b.eps;
        b.lambda[p] = -C / denom;
        c2 = (double)C * (double)C;
    }
    const double s = block_sum(c2);
    if (threadIdx.x == 0) b.part[(size_t)b.it * b.nblocks + blockIdx.x] = s;
}

// Folds iterate b.it's block sums (one block of 1024: strided sums in block
// order, then the waves in order) and decides whether it has converged.
__global__ __launch_bounds__(1024) void k_pbf_check(PbfArgs b) {
    __shared__ double part[16];
    if (b.ctl->done) return;
    double v = 0.0;
    for (int k = threadIdx.x; k < b.nblocks; k += 1024) v += b.part[(size_t)b.it * b.nblocks + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < 16; ++k) s += part[k];
        b.res[b.it] = s;
        if (s <= b.tol2n) {
            b.ctl->done = 1;
            b.ctl->final_it = b.it;
        }
    }
}

// The update of iterate b.it into the next iterate (nx, ny).
__global__ __launch_bounds__(256) void k_pbf_delta(PbfArgs b) {
    if (b.ctl->done) return;
    __shared__ Staged st;
    stage_windows<true>(b, st);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.n) return;
    const float lp = b.lambda[p];
    float ax = 0.0f, ay = 0.0f;
    for_listed(b, st, p, [&](uint32_t q, uint32_t slot, float dx, float dy, float d2) {
        const float w = spiky_w(b, d2);
        const bool lds = slot < (uint32_t)kStage;
        const float s = lp + (lds ? st.l[lds ? slot : 0] : b.lambda[q]);
        ax += s * (w * dx);
        ay += s * (w * dy);
    });
    b.nx[p] = wrap_pos(b.xs[p] + -(ax / b.rho0), b.W);
    b.ny[p] = wrap_pos(b.ys[p] + -(ay / b.rho0), b.H);
}

// Ends the solve on the caller's buffers: the final iterate f lives in the
// scratch pair when f is odd.
__global__ void k_pbf_finish(const PbfCtl *ctl, int iters, const float *sx, const float *sy,
                             float *xs, float *ys, int *iters_used, int n) {
    const int f = ctl->done ? ctl->final_it : iters;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && iters_used) *iters_used = f;
    if (p < n && (f & 1)) {
        xs[p] = sx[p];
        ys[p] = sy[p];
    }
}

__global__ void k_pbf_init(PbfCtl *ctl, double *res, int nres) {
    if (threadIdx.x == 0) {
        ctl->done = 0;
        ctl->final_it = 0;
    }
    for (int k = threadIdx.x; k < nres; k += blockDim.x) res[k] = -1.0;  // not computed
}

int grid_for(int n) { return (n + 255) / 256; }

// The digit plan for keys below ncells: the fewest passes, 8-bit digits
// where they suffice for that many passes, else 10-bit.
struct RadixPlan {
    int passes, bits;
};
RadixPlan radix_plan(uint32_t ncells) {
    int kb = 0;
    while (kb < 32 && (ncells - 1) >> kb) ++kb;
    if (kb <= 8) return {1, 8};
    if (kb <= 16) return {2, 8};
    if (kb <= 20) return {2, 10};
    if (kb <= 24) return {3, 8};
    return {3, 10};
}

// Workspace layout (words): sorted keys (2 buffers), values spare buffer,
// digit histogram and its row scans (2^10 * nblocks each), digit row totals
// (2^10), scan block sums.
struct SortWs {
    uint32_t *k0, *k1, *v1, *hist, *hscan, *tot, *sums;
};
size_t hist_words(int n) {
    const size_t nb = (size_t)nblocks_for((size_t)n);
    return ((size_t)1 << kMaxRadixBits) * (nb ? nb : 1);
}
size_t sort_ws_words(int n, size_t ncells) {
    const size_t hist = hist_words(n);
    size_t sums = std::max<size_t>(nblocks_for(hist), nblocks_for((size_t)n));
    sums = std::max<size_t>(sums, nblocks_for(ncells));
    return 3 * (size_t)n + 2 * hist + ((size_t)1 << kMaxRadixBits) + sums + 64;
}
SortWs sort_ws(void *ws, int n) {  // ws: 4-byte aligned
    uint32_t *p = static_cast<uint32_t *>(ws);
    SortWs s;
    s.k0 = p;
    s.k1 = p + n;
    s.v1 = p + 2 * (size_t)n;
    s.hist = p + 3 * (size_t)n;
    s.hscan = s.hist + hist_words(n);
    s.tot = s.hscan + hist_words(n);
    s.sums = s.tot + ((size_t)1 << kMaxRadixBits);
    return s;
}

bool valid_grid(const pfs_ngrid *g) {
    return g && g->n >= 1 && g->gx >= 3 && g->gy >= 3 && g->domain_w > 0 && g->domain_h > 0 &&
           g->radius > 0 && (int64_t)g->gx * g->gy <= ((int64_t)1 << 30) &&
           g->radius <= g->domain_w / g->gx && g->radius <= g->domain_h / g->gy;
}

NArgs nargs(const pfs_ngrid *g, const float *xs, const float *ys,
            const uint32_t *order, const uint32_t *sk, const uint32_t *start, const uint32_t *end) {
    NArgs a{};
    a.xs = xs;
    a.ys = ys;
    a.order = order;
    a.sk = sk;
    a.start = start;
    a.end = end;
    a.n = g->n;
    a.gx = g->gx;
    a.gy = g->gy;
    a.W = g->domain_w;
    a.H = g->domain_h;
    a.half_w = g->domain_w * 0.5f;
    a.half_h = g->domain_h * 0.5f;
    a.r2 = g->radius * g->radius;
    return a;
}

// PBF workspace: the control block, the block sums of C^2 of every iterate,
// and the scratch position pair.
struct PbfWs {
    PbfCtl *ctl;
    double *part;
    float *sx, *sy;
};
size_t pbf_part_bytes(int n, int max_iters) {
    return ((size_t)(max_iters + 1) * grid_for(n) * sizeof(double) + 255) & ~(size_t)255;
}
size_t pbf_ws_bytes(int n, int max_iters) {
    return 256 + pbf_part_bytes(n, max_iters) + 2 * (((size_t)n * sizeof(float) + 255) & ~(size_t)255);
}
PbfWs pbf_ws(void *ws, int n, int max_iters) {
    char *p = static_cast<char *>(ws);
    PbfWs w;
    w.ctl = reinterpret_cast<PbfCtl *>(p);
    w.part = reinterpret_cast<double *>(p + 256);
    w.sx = reinterpret_cast<float *>(p + 256 + pbf_part_bytes(n, max_iters));
    w.sy = w.sx + ((((size_t)n * sizeof(float) + 255) & ~(size_t)255) / sizeof(float));
    return w;
}

// Offsets and *total from counts, then the lists (checked against the
// capacity first, which waits for *total).
int lists_from_counts(const pfs_ngrid *g, const NArgs &a, const uint32_t *counts,
                      uint32_t *offsets, uint32_t *nbr, size_t nbr_capacity, uint32_t *total,
                      const SortWs &w, hipStream_t s) {
    PFS_TRY(scan_exclusive(counts, offsets, g->n, total, w.sums, s));
    uint32_t h_total = 0;
    PFS_HIP_TRY(hipMemcpyAsync(&h_total, total, sizeof h_total, hipMemcpyDeviceToHost, s));
    PFS_HIP_TRY(hipStreamSynchronize(s));
    if (h_total > nbr_capacity)
        return set_error(PFS_ERR_INVALID, "ngrid: neighbour list capacity " +
                                              std::to_string(nbr_capacity) + " < " +
                                              std::to_string(h_total));
    if (a.order)
        hipLaunchKernelGGL(k_nbr_fill, dim3(grid_for(g->n)), dim3(256), 0, s, a, offsets, nbr,
                           nbr_capacity);
    else
        hipLaunchKernelGGL(k_nbr_fill_lds, dim3(grid_for(g->n)), dim3(256), 0, s, a, offsets,
                           counts, nbr, nbr_capacity);
    return check_launch("neighbour lists");
}

}  // namespace
}  // namespace pfs

using namespace pfs;

extern "C" {

int pfs_ngrid_init(pfs_ngrid *g, int n, float domain_w, float domain_h, float radius) {
    if (!g || n < 1 || !(domain_w > 0) || !(domain_h > 0) || !(radius > 0))
        return set_error(PFS_ERR_INVALID, "ngrid: need n >= 1 and positive sizes");
    g->n = n;
    g->domain_w = domain_w;
    g->domain_h = domain_h;
    g->radius = radius;
    const double gx = std::floor((double)domain_w / radius), gy = std::floor((double)domain_h / radius);
    if (gx < 3 || gy < 3 || gx * gy > (double)(1 << 30))
        return set_error(PFS_ERR_INVALID, "ngrid: the domain must span 3 to 2^30 cells of the radius");
    g->gx = (int)gx;
    g->gy = (int)gy;
    // Cells at least `radius` wide, so neighbours within it are in the 3x3 block.
    while (g->radius > g->domain_w / g->gx) --g->gx;
    while (g->radius > g->domain_h / g->gy) --g->gy;
    if (g->gx < 3 || g->gy < 3) return set_error(PFS_ERR_INVALID, "ngrid: fewer than 3 cells");
    g->kx = (float)g->gx / domain_w;
    g->ky = (float)g->gy / domain_h;
    return PFS_OK;
}

size_t pfs_ngrid_workspace_bytes(const pfs_ngrid *g) {
    if (!g || g->n < 1) return 0;
    return sort_ws_words(g->n, (size_t)g->gx * g->gy) * sizeof(uint32_t);
}

int pfs_scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *total, void *ws,
                           size_t ws_bytes, void *stream) {
    if (n < 0 || (n > 0 && (!in || !out))) return set_error(PFS_ERR_INVALID, "scan: bad arguments");
    if (!ws || ws_bytes < (size_t)(nblocks_for((size_t)n) + 1) * sizeof(uint32_t))
        return set_error(PFS_ERR_INVALID, "scan: workspace too small");
    return scan_exclusive(in, out, n, total, static_cast<uint32_t *>(ws), (hipStream_t)stream);
}

int pfs_gather_u32(const uint32_t *index, const uint32_t *in, uint32_t *out, int n, void *stream) {
    if (n < 0 || (n > 0 && (!index || !in || !out)) || in == out)
        return set_error(PFS_ERR_INVALID, "gather: bad arguments");
    if (n == 0) return PFS_OK;
    hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, index, in,
                       out, n);
    return check_launch("gather");
}

int pfs_particles_advect(float *x, float *y, int n, const pfs_field *vp, float dt, void *stream) {
    PFS_TRY(check_field(vp, "vp", 2));
    if (n < 0 || (n > 0 && (!x || !y))) return set_error(PFS_ERR_INVALID, "advect: bad particles");
    if (n == 0) return PFS_OK;
    hipLaunchKernelGGL(k_tracer_advect, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y,
                       n, vp->data, vp->x, vp->y, vp->z, dt, (float)vp->x, (float)vp->y,
                       1.0f / (float)vp->x, 1.0f / (float)vp->y, (vp->x & (vp->x - 1)) == 0,
                       (vp->y & (vp->y - 1)) == 0);
    return check_launch("particle advect");
}

int pfs_ngrid_build(const pfs_ngrid *g, const float *x, const float *y, uint32_t *keys,
                    uint32_t *order, uint32_t *cell_start, uint32_t *cell_end, float *xs, float *ys,
                    void *ws, size_t ws_bytes, void *stream) {
    if (!valid_grid(g)) return set_error(PFS_ERR_INVALID, "ngrid: not initialised by pfs_ngrid_init");
    if (!x || !y || !keys || !order || !cell_start || !cell_end || !xs || !ys)
        return set_error(PFS_ERR_INVALID, "ngrid: NULL buffer");
    if (!ws || ws_bytes < pfs_ngrid_workspace_bytes(g) || ((uintptr_t)ws & 3))
        return set_error(PFS_ERR_INVALID,
                         "ngrid: workspace not 4-byte aligned or smaller than pfs_ngrid_workspace_bytes()");
    hipStream_t s = (hipStream_t)stream;
    const int n = g->n, nb = nblocks_for((size_t)n);
    const uint32_t ncells = (uint32_t)g->gx * (uint32_t)g->gy;
    SortWs w = sort_ws(ws, n);

    // Buffers alternate so that the last pass writes the keys to k0 (kept for
    // the neighbour passes) and the particle ids to `order`.  The first pass
    // makes the keys (into `keys`) and takes the particle ids as its values;
    // the last gathers the positions into sorted order.
    const RadixPlan plan = radix_plan(ncells);
    const int passes = plan.passes;
    uint32_t *kb[2] = {passes % 2 == 0 ? w.k0 : w.k1, passes % 2 == 0 ? w.k1 : w.k0};
    uint32_t *vb[2] = {passes % 2 == 0 ? order : w.v1, passes % 2 == 0 ? w.v1 : order};
    const CellKeyArgs ck{x, y, g->kx, g->ky, g->gx, g->gy};
    const uint32_t *kin = keys;
    for (int p = 0; p < passes; ++p) {
        const int shift = plan.bits * p;
        const bool last = p + 1 == passes;
        const uint32_t *vin = p == 0 ? nullptr : vb[p & 1];
        float *oxs = last ? xs : nullptr, *oys = last ? ys : nullptr;
        if (plan.bits == 8) {
            if (p == 0)
                hipLaunchKernelGGL((k_radix_hist<8, true>), dim3(nb), dim3(kThreads), 0, s, nullptr,
                                   ck, keys, n, shift, w.hist, nb);
            else
                hipLaunchKernelGGL((k_radix_hist<8, false>), dim3(nb), dim3(kThreads), 0, s, kin,
                                   ck, nullptr, n, shift, w.hist, nb);
            hipLaunchKernelGGL(k_digit_rows, dim3(1 << 8), dim3(kThreads), 0, s, w.hist, w.hscan,
                               w.tot, nb);
            hipLaunchKernelGGL(k_radix_scatter<8>, dim3(nb), dim3(kThreads), 0, s, kin, vin,
                               kb[(p + 1) & 1], vb[(p + 1) & 1], n, shift, w.hscan, w.tot, nb, x, y,
                               oxs, oys);
        } else {
            if (p == 0)
                hipLaunchKernelGGL((k_radix_hist<10, true>), dim3(nb), dim3(kThreads), 0, s, nullptr,
                                   ck, keys, n, shift, w.hist, nb);
            else
                hipLaunchKernelGGL((k_radix_hist<10, false>), dim3(nb), dim3(kThreads), 0, s, kin,
                                   ck, nullptr, n, shift, w.hist, nb);
            hipLaunchKernelGGL(k_digit_rows, dim3(1 << 10), dim3(kThreads), 0, s, w.hist, w.hscan,
                               w.tot, nb);
            hipLaunchKernelGGL(k_radix_scatter<10>, dim3(nb), dim3(kThreads), 0, s, kin, vin,
                               kb[(p + 1) & 1], vb[(p + 1) & 1], n, shift, w.hscan, w.tot, nb, x, y,
                               oxs, oys);
        }
        kin = kb[(p + 1) & 1];
    }
    // Cell tables: the lower bound of every cell (empty ones included) in the
    // sorted keys, left in k0 by the last pass.
    const unsigned cblocks = (unsigned)(((size_t)ncells + 255) / 256);
    hipLaunchKernelGGL(k_cell_tables, dim3(cblocks), dim3(256), 0, s, w.k0, n, cell_start, cell_end,
                       ncells);
    return check_launch("neighbour grid");
}

int pfs_ngrid_neighbors(const pfs_ngrid *g, const float *xs, const float *ys, const uint32_t *order,
                        const uint32_t *cell_start, const uint32_t *cell_end, uint32_t *counts,
                        uint32_t *offsets, uint32_t *nbr, size_t nbr_capacity, uint32_t *total,
                        void *ws, size_t ws_bytes, void *stream) {
    if (!valid_grid(g)) return set_error(PFS_ERR_INVALID, "ngrid: not initialised by pfs_ngrid_init");
    if (!xs || !ys || !cell_start || !cell_end || !counts || !offsets || !total)
        return set_error(PFS_ERR_INVALID, "ngrid: NULL buffer");
    if (!ws || ws_bytes < pfs_ngrid_workspace_bytes(g) || ((uintptr_t)ws & 3))
        return set_error(PFS_ERR_INVALID,
                         "ngrid: workspace not 4-byte aligned or smaller than pfs_ngrid_workspace_bytes()");
    hipStream_t s = (hipStream_t)stream;
    SortWs w = sort_ws(ws, g->n);
    const NArgs a = nargs(g, xs, ys, order, w.k0, cell_start, cell_end);
    hipLaunchKernelGGL(k_nbr_count, dim3(grid_for(g->n)), dim3(256), 0, s, a, counts);
    if (!nbr) return scan_exclusive(counts, offsets, g->n, total, w.sums, s);  // no lists
    return lists_from_counts(g, a, counts, offsets, nbr, nbr_capacity, total, w, s);
}

int pfs_ngrid_lists(const pfs_ngrid *g, const float *xs, const float *ys, const uint32_t *order,
                    const uint32_t *cell_start, const uint32_t *cell_end, const uint32_t *counts,
                    uint32_t *offsets, uint32_t *nbr, size_t nbr_capacity, uint32_t *total,
                    void *ws, size_t ws_bytes, void *stream) {
    if (!valid_grid(g)) return set_error(PFS_ERR_INVALID, "ngrid: not initialised by pfs_ngrid_init");
    if (!xs || !ys || !cell_start || !cell_end || !counts || !offsets || !nbr || !total)
        return set_error(PFS_ERR_INVALID, "ngrid: NULL buffer");
    if (!ws || ws_bytes < pfs_ngrid_workspace_bytes(g) || ((uintptr_t)ws & 3))
        return set_error(PFS_ERR_INVALID,
                         "ngrid: workspace not 4-byte aligned or smaller than pfs_ngrid_workspace_bytes()");
    SortWs w = sort_ws(ws, g->n);
    const NArgs a = nargs(g, xs, ys, order, w.k0, cell_start, cell_end);
    return lists_from_counts(g, a, counts, offsets, nbr, nbr_capacity, total, w,
                             (hipStream_t)stream);
}


size_t pfs_pbf_workspace_bytes(const pfs_ngrid *g, int max_iters) {
    if (!g || g->n < 1 || max_iters < 0 || max_iters > PFS_PBF_MAX_ITERS) return 0;
    return pbf_ws_bytes(g->n, max_iters);
}

int pfs_pbf_solve(const pfs_ngrid *g, float *xs, float *ys, const uint32_t *cell_start,
                  const uint32_t *cell_end, const uint32_t *nbr_offsets, const uint32_t *nbr_counts,
                  const uint32_t *nbr, float mass, float rho0, float eps, int max_iters, double tol,
                  float *lambda, double *residuals, int *iters_used, void *ngrid_ws,
                  size_t ngrid_ws_bytes, void *ws, size_t ws_bytes, void *stream) {
    if (!valid_grid(g)) return set_error(PFS_ERR_INVALID, "ngrid: not initialised by pfs_ngrid_init");
    if (!xs || !ys || !cell_start || !cell_end || !nbr_offsets || !nbr_counts || !nbr || !lambda ||
        !residuals)
        return set_error(PFS_ERR_INVALID, "pbf: NULL buffer");
    if (!ngrid_ws || ngrid_ws_bytes < pfs_ngrid_workspace_bytes(g) || ((uintptr_t)ngrid_ws & 3))
        return set_error(PFS_ERR_INVALID, "pbf: neighbour-grid workspace too small or misaligned");
    if (!(rho0 > 0.0f) || !(eps > 0.0f) || !(mass > 0.0f) || !(tol >= 0.0) || max_iters < 0 ||
        max_iters > PFS_PBF_MAX_ITERS)
        return set_error(PFS_ERR_INVALID, "pbf: need mass, rho0, eps > 0, tol >= 0, 0 <= max_iters <= " +
                                              std::to_string(PFS_PBF_MAX_ITERS));
    if (!ws || ws_bytes < pbf_ws_bytes(g->n, max_iters) || ((uintptr_t)ws & 255))
        return set_error(PFS_ERR_INVALID,
                         "pbf: workspace not 256-byte aligned or smaller than pfs_pbf_workspace_bytes()");
    hipStream_t s = (hipStream_t)stream;
    const int n = g->n, nb = grid_for(n);
    const PbfWs w = pbf_ws(ws, n, max_iters);
    const double h = g->radius;
    PbfArgs b{};
    b.noff = nbr_offsets;
    b.ncnt = nbr_counts;
    b.nbr = nbr;
    b.sk = sort_ws(ngrid_ws, n).k0;
    b.cstart = cell_start;
    b.cend = cell_end;
    b.gx = g->gx;
    b.gy = g->gy;
    b.n = n;
    b.W = g->domain_w;
    b.H = g->domain_h;
    b.half_w = g->domain_w * 0.5f;
    b.half_h = g->domain_h * 0.5f;
    b.h = g->radius;
    b.h2 = g->radius * g->radius;
    const float coef = (float)(4.0 / (3.14159265358979323846 * h * h * h * h * h * h * h * h));
    b.mc = mass * coef;
    b.gc = (float)(-30.0 / (3.14159265358979323846 * h * h * h * h * h));
    b.rho0 = rho0;
    b.r0sq = rho0 * rho0;
    b.eps = eps;
    b.lambda = lambda;
    b.part = w.part;
    b.res = residuals;
    b.ctl = w.ctl;
    b.nblocks = nb;
    b.tol2n = tol * tol * (double)n;
    float *bx[2] = {xs, w.sx}, *by[2] = {ys, w.sy};
    hipLaunchKernelGGL(k_pbf_init, dim3(1), dim3(256), 0, s, w.ctl, residuals, max_iters + 1);
    for (int it = 0; it <= max_iters; ++it) {
        b.it = it;
        b.xs = bx[it & 1];
        b.ys = by[it & 1];
        b.nx = bx[(it + 1) & 1];
        b.ny = by[(it + 1) & 1];
        hipLaunchKernelGGL(k_pbf_lambda, dim3(nb), dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_pbf_check, dim3(1), dim3(1024), 0, s, b);
        if (it < max_iters) hipLaunchKernelGGL(k_pbf_delta, dim3(nb), dim3(256), 0, s, b);
    }
    hipLaunchKernelGGL(k_pbf_finish, dim3(nb), dim3(256), 0, s, w.ctl, max_iters, w.sx, w.sy, xs, ys,
                       iters_used, n);
    return check_launch("pbf solve");
}

int pfs_ngrid_density(const pfs_ngrid *g, const float *xs, const float *ys, const uint32_t *order,
                      const uint32_t *cell_start, const uint32_t *cell_end, float mass, float *rho,
                      uint32_t *counts, void *ws, size_t ws_bytes, void *stream) {
    if (!valid_grid(g)) return set_error(PFS_ERR_INVALID, "ngrid: not initialised by pfs_ngrid_init");
    if (!xs || !ys || !cell_start || !cell_end || !rho)
        return set_error(PFS_ERR_INVALID, "ngrid: NULL buffer");
    if (!ws || ws_bytes < pfs_ngrid_workspace_bytes(g) || ((uintptr_t)ws & 3))
        return set_error(PFS_ERR_INVALID,
                         "ngrid: workspace not 4-byte aligned or smaller than pfs_ngrid_workspace_bytes()");
    SortWs w = sort_ws(ws, g->n);
    hipStream_t s = (hipStream_t)stream;
    const NArgs a = nargs(g, xs, ys, order, w.k0, cell_start, cell_end);
    // 2D poly6: 4 / (pi r^8), in float from the double constant, times the mass.
    const double r = g->radius;
    const float coef = (float)(4.0 / (3.14159265358979323846 * r * r * r * r * r * r * r * r));
    const float mc = mass * coef;
    hipLaunchKernelGGL(k_density, dim3(grid_for(g->n)), dim3(256), 0, s, a, mc, rho, counts);
    return check_launch("density");
}

}  // extern "C"
