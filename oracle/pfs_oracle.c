/*
 * CPU oracle for the grid fluid step — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain C restatement of the reference's CPU solver, src/fluid.cpp, with
 * the buffer ping-pong, the stale channels and the even/odd Jacobi-count
 * behaviour of that file reproduced exactly.  It is the checker the HIP
 * library is compared against, bit for bit.  Only tests/,
 * __graft_entry__.smoke() and bench.py (the CPU baseline, and the checker of
 * its legs' results outside the timed regions) may load it; the product
 * library never links or calls it.
 *
 * Two instantiations:
 *   orc_*    float, the reference's own precision (parity oracle)
 *   orc64_*  double twin of the same algorithm (the error yardstick of the
 *            precision form; not a reference behaviour)
 *
 * The reference's compile-time Jacobi count NUM_JACOBI_ITERS
 * (includes/fluid.hpp:11) is the runtime argument `iters` here.
 *
 * How it is pinned: tests/test_oracle.py checks it bit-exactly against the
 * golden vectors in tests/golden/ (written by the reference's own fluid.cpp,
 * compiled by oracle/Makefile) and against closed forms: zero velocity and
 * whole-cell shifts for advection, the Fourier-mode eigenvalues of both
 * Jacobi sweeps, constant-field invariants of divergence and gradient.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { int x, y, z; float *data; } orc_field;
typedef struct { int x, y, z; double *data; } orc64_field;

#define R float
#define FIELD orc_field
#define FMOD fmodf
#define NAME(s) orc_##s
#include "pfs_oracle_impl.h"
#undef R
#undef FIELD
#undef FMOD
#undef NAME

#define R double
#define FIELD orc64_field
#define FMOD fmod
#define NAME(s) orc64_##s
#include "pfs_oracle_impl.h"
#undef R
#undef FIELD
#undef FMOD
#undef NAME

/* Thread count for the row-parallel loops (1 = the reference's single thread). */
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * Residual of the pressure Poisson system solved by compute_pressure's sweep
 * p' = (pL+pR+pT+pB+b)/4 (src/fluid.cpp:249-255): at its fixed point
 * r = b + pL + pR + pT + pB - 4p = 0.  Returns sum r^2 and sum b^2 in double,
 * reading p from channel 2 and b from channel 3 of `vp`.  This quantity has
 * no counterpart in the reference; it is the convergence measure of the
 * precision form.
 */
void orc_pressure_residual(const orc_field *vp, double *sum_r2, double *sum_b2) {
    const int w = vp->x, h = vp->y, c = vp->z;
    const float *d = vp->data;
    double r2 = 0.0, b2 = 0.0;
    for (int j = 0; j < h; j++) {
        for (int i = 0; i < w; i++) {
            const int im = (i - 1 + w) % w, ip = (i + 1) % w;
            const int jm = (j - 1 + h) % h, jp = (j + 1) % h;
            const double p = d[(j * w + i) * c + 2];
            const double b = d[(j * w + i) * c + 3];
            const double s = (double)d[(j * w + im) * c + 2] + (double)d[(j * w + ip) * c + 2] +
                             (double)d[(jm * w + i) * c + 2] + (double)d[(jp * w + i) * c + 2];
            const double r = b + s - 4.0 * p;
            r2 += r * r;
            b2 += b * b;
        }
    }
    *sum_r2 = r2;
    *sum_b2 = b2;
}
