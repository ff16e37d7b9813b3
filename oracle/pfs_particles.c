/* CPU restatement of the particle layer (tracers and neighbour grid) --
 * TEST INFRASTRUCTURE ONLY.
 *
 * The reference has no particles; this is the oracle for the library's
 * pfs_particles_advect / pfs_ngrid_* (probabilistic_fluid_simulation_amd/
 * csrc/pfs_particles.hip), written plainly: the same float operations in
 * the same order (compiled with -ffp-contract=off), a counting sort where
 * the GPU radix-sorts, and the same neighbour traversal.  Its integer
 * outputs are checked against numpy (stable argsort, brute-force O(n^2)
 * neighbour sets) in tests/test_particles_oracle.py. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static float bilerp(float sx, float sy, float f00, float f10, float f01, float f11) {
    /* src/fluid.cpp:62-66, left to right */
    return (1 - sx) * (1 - sy) * f00 + sx * (1 - sy) * f10 + (1 - sx) * sy * f01 + sx * sy * f11;
}

static float wrap_pos(float x, float n) {
    if (x >= 0.0f && x < n) return x;
    x = fmodf(x, n);
    if (x < 0.0f) x += n;
    return x >= n ? 0.0f : x;
}

static float div_n(float x, float n, float inv_n, int pow2) { return pow2 ? x * inv_n : x / n; }

void orcp_advect(float *x, float *y, int n, const float *vp, int w, int h, int vc, float dt) {
    const float fw = (float)w, fh = (float)h, inv_w = 1.0f / fw, inv_h = 1.0f / fh;
    const int pw = (w & (w - 1)) == 0, ph = (h & (h - 1)) == 0;
    for (int i = 0; i < n; ++i) {
        const float px = x[i], py = y[i];
        int i0 = (int)px, j0 = (int)py;
        i0 = i0 < 0 ? 0 : (i0 > w - 1 ? w - 1 : i0);
        j0 = j0 < 0 ? 0 : (j0 > h - 1 ? h - 1 : j0);
        const int i1 = i0 + 1 == w ? 0 : i0 + 1, j1 = j0 + 1 == h ? 0 : j0 + 1;
        const float sx = px - (float)i0, sy = py - (float)j0;
        const float *f00 = vp + ((size_t)j0 * w + i0) * vc, *f10 = vp + ((size_t)j0 * w + i1) * vc;
        const float *f01 = vp + ((size_t)j1 * w + i0) * vc, *f11 = vp + ((size_t)j1 * w + i1) * vc;
        const float u = bilerp(sx, sy, f00[0], f10[0], f01[0], f11[0]);
        const float v = bilerp(sx, sy, f00[1], f10[1], f01[1], f11[1]);
        x[i] = wrap_pos(px + div_n(dt * u, fw, inv_w, pw), fw);
        y[i] = wrap_pos(py + div_n(dt * v, fh, inv_h, ph), fh);
    }
}

static int cell_of(float x, float k, int g) {
    const int c = (int)(x * k);
    return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

/* keys, order (stable by key), cell bounds, sorted positions. */
int orcp_ngrid_build(const float *x, const float *y, int n, float kx, float ky, int gx, int gy,
                     uint32_t *keys, uint32_t *order, uint32_t *start, uint32_t *end, float *xs,
                     float *ys) {
    const size_t nc = (size_t)gx * gy;
    uint32_t *count = calloc(nc + 1, sizeof(uint32_t));
    if (!count) return -1;
    for (int i = 0; i < n; ++i) {
        keys[i] = (uint32_t)(cell_of(y[i], ky, gy) * gx + cell_of(x[i], kx, gx));
        count[keys[i] + 1]++;
    }
    for (size_t c = 0; c < nc; ++c) count[c + 1] += count[c];  /* count[c] = start of c */
    for (size_t c = 0; c < nc; ++c) {  /* every cell: [lower bound, upper bound) */
        start[c] = count[c];
        end[c] = count[c + 1];
    }
    for (int i = 0; i < n; ++i) order[count[keys[i]]++] = (uint32_t)i;
    for (int p = 0; p < n; ++p) {
        xs[p] = x[order[p]];
        ys[p] = y[order[p]];
    }
    free(count);
    return 0;
}

/* x_q - x_p on a periodic axis of length L (half = L/2): the nearer image,
 * with the far operand moved by L before the subtraction (exact for a
 * position in (L/2, L)), so that only nearby values are subtracted. */
static float periodic_diff(float xq, float xp, float L, float half) {
    const float d = xq - xp;
    return d > half ? (xq - L) - xp : (d < -half ? xq - (xp - L) : d);
}

/* Neighbours of every particle (ids in traversal order), counts by particle
 * id; nbr may be NULL (counts only); order NULL: ids are sorted positions.
 * Returns the total. */
long long orcp_ngrid_neighbors(const float *xs, const float *ys, const uint32_t *order, int n,
                               const uint32_t *keys_sorted, const uint32_t *start,
                               const uint32_t *end, int gx, int gy, float W, float H, float r,
                               uint32_t *counts, const uint32_t *offsets, uint32_t *nbr) {
    const float half_w = W * 0.5f, half_h = H * 0.5f, r2 = r * r;
    long long total = 0;
    for (int p = 0; p < n; ++p) {
        const uint32_t c = keys_sorted[p];
        const int cx = (int)(c % (uint32_t)gx), cy = (int)(c / (uint32_t)gx);
        const float px = xs[p], py = ys[p];
        const uint32_t id = order ? order[p] : (uint32_t)p;
        uint32_t k = 0, o = (nbr && offsets) ? offsets[id] : 0;
        for (int oy = -1; oy <= 1; ++oy) {
            const int ny = (cy + oy + gy) % gy;
            for (int ox = -1; ox <= 1; ++ox) {
                const int nx = (cx + ox + gx) % gx;
                const uint32_t cc = (uint32_t)(ny * gx + nx);
                for (uint32_t q = start[cc]; q < end[cc]; ++q) {
                    if ((int)q == p) continue;
                    const float dx = periodic_diff(xs[q], px, W, half_w);
                    const float dy = periodic_diff(ys[q], py, H, half_h);
                    const float d2 = dx * dx + dy * dy;
                    if (d2 < r2) {
                        if (nbr && offsets) nbr[o + k] = order ? order[q] : q;
                        ++k;
                    }
                }
            }
        }
        counts[id] = k;
        total += k;
    }
    return total;
}

/* SPH density (2D poly6): rho = (mass * coef) * S, S the float sum in
 * traversal order of (r^2 - d2)^3 over the candidates with d2 < r^2, the
 * particle itself included where the traversal meets it. */
void orcp_ngrid_density(const float *xs, const float *ys, const uint32_t *order, int n,
                        const uint32_t *keys_sorted, const uint32_t *start, const uint32_t *end,
                        int gx, int gy, float W, float H, float r, float mass, float *rho) {
    const float half_w = W * 0.5f, half_h = H * 0.5f, r2 = r * r;
    const double rd = r;
    const float coef = (float)(4.0 / (3.14159265358979323846 * rd * rd * rd * rd * rd * rd * rd * rd));
    const float mc = mass * coef;
    for (int p = 0; p < n; ++p) {
        float sum = 0.0f;
        const uint32_t c = keys_sorted[p];
        const int cx = (int)(c % (uint32_t)gx), cy = (int)(c / (uint32_t)gx);
        const float px = xs[p], py = ys[p];
        for (int oy = -1; oy <= 1; ++oy) {
            const int ny = (cy + oy + gy) % gy;
            for (int ox = -1; ox <= 1; ++ox) {
                const int nx = (cx + ox + gx) % gx;
                const uint32_t cc = (uint32_t)(ny * gx + nx);
                for (uint32_t q = start[cc]; q < end[cc]; ++q) {
                    const float dx = periodic_diff(xs[q], px, W, half_w);
                    const float dy = periodic_diff(ys[q], py, H, half_h);
                    const float d2 = dx * dx + dy * dy;
                    if (d2 < r2) {
                        const float t = r2 - d2;
                        sum += t * t * t;
                    }
                }
            }
        }
        rho[order ? order[p] : (uint32_t)p] = mc * sum;
    }
}

/* Position-based fluids (pfs_pbf_solve; Macklin & Mueller 2013), the same
 * float operations in the same order: per iterate k, lambda of every
 * particle from iterate k (its own term, then its neighbour list in order),
 * sum C^2 in FP64 (sequential here, by blocks in the library: equal to
 * rounding), the convergence test, then every update from iterate k.  The
 * lists (by sorted position) are those of the starting positions. */
static float spiky_w(float h, float h2, float gc, float d2) {
    if (!(d2 > 0.0f) || !(d2 < h2)) return 0.0f;
    const float r = sqrtf(d2);
    const float t = h - r;
    return ((gc * t) * t) / r;
}

int orcp_pbf_solve(float *xs, float *ys, int n, const uint32_t *offsets, const uint32_t *counts,
                   const uint32_t *nbr, float W, float H, float r, float mass, float rho0, float eps,
                   int max_iters, double tol, float *lambda, double *res, int *iters_used) {
    const float half_w = W * 0.5f, half_h = H * 0.5f, h = r, h2 = r * r;
    const double hd = r;
    const float coef = (float)(4.0 / (3.14159265358979323846 * hd * hd * hd * hd * hd * hd * hd * hd));
    const float mc = mass * coef;
    const float gc = (float)(-30.0 / (3.14159265358979323846 * hd * hd * hd * hd * hd));
    const float r0sq = rho0 * rho0;
    float *nx = malloc(sizeof(float) * (size_t)n), *ny = malloc(sizeof(float) * (size_t)n);
    if (!nx || !ny) {
        free(nx);
        free(ny);
        return -1;
    }
    for (int k = 0; k <= max_iters; ++k) res[k] = -1.0;
    int it = 0;
    for (;; ++it) {
        double c2 = 0.0;
        for (int p = 0; p < n; ++p) {
            float S = h2 * h2 * h2, gsx = 0.0f, gsy = 0.0f, g2 = 0.0f;
            for (uint32_t k = 0; k < counts[p]; ++k) {
                const uint32_t q = nbr[offsets[p] + k];
                const float dx = periodic_diff(xs[q], xs[p], W, half_w);
                const float dy = periodic_diff(ys[q], ys[p], H, half_h);
                const float d2 = dx * dx + dy * dy;
                if (d2 < h2) {
                    const float t2 = h2 - d2;
                    S += t2 * t2 * t2;
                }
                const float w = spiky_w(h, h2, gc, d2);
                if (w != 0.0f) {
                    const float vx = w * dx, vy = w * dy;
                    gsx += vx;
                    gsy += vy;
                    g2 += vx * vx + vy * vy;
                }
            }
            const float rho = mc * S;
            const float C = rho / rho0 - 1.0f;
            const float denom = (gsx * gsx + gsy * gsy + g2) / r0sq + eps;
            lambda[p] = -C / denom;
            c2 += (double)C * (double)C;
        }
        res[it] = c2;
        if (c2 <= tol * tol * (double)n || it == max_iters) break;
        for (int p = 0; p < n; ++p) {
            const float lp = lambda[p];
            float ax = 0.0f, ay = 0.0f;
            for (uint32_t k = 0; k < counts[p]; ++k) {
                const uint32_t q = nbr[offsets[p] + k];
                const float dx = periodic_diff(xs[q], xs[p], W, half_w);
                const float dy = periodic_diff(ys[q], ys[p], H, half_h);
                const float w = spiky_w(h, h2, gc, dx * dx + dy * dy);
                if (w != 0.0f) {
                    const float s = lp + lambda[q];
                    ax += s * (w * dx);
                    ay += s * (w * dy);
                }
            }
            nx[p] = wrap_pos(xs[p] + -(ax / rho0), W);
            ny[p] = wrap_pos(ys[p] + -(ay / rho0), H);
        }
        memcpy(xs, nx, sizeof(float) * (size_t)n);
        memcpy(ys, ny, sizeof(float) * (size_t)n);
    }
    *iters_used = it;
    free(nx);
    free(ny);
    return 0;
}
