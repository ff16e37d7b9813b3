/*
 * Body of the CPU oracle, instantiated once per scalar type by pfs_oracle.c.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain restatement of the reference's
 * CPU solver (src/fluid.cpp) used as the checker for the HIP library.  Only
 * tests/, __graft_entry__.smoke() and bench.py (the CPU baseline, and the
 * checker of its legs' results) may load it.
 *
 * The including file defines:
 *   R        scalar type (float or double)
 *   FIELD    field struct type {int x, y, z; R *data;}
 *   FMOD     fmodf or fmod
 *   NAME(s)  symbol prefixer
 *
 * Every expression keeps the reference's operation order exactly (left-to-
 * right sums, products grouped as C parses them, IEEE division).  The file is
 * compiled with -ffp-contract=off so no multiply-add is fused.
 */

/* Row-major interleaved index, src/fluid.cpp:15-17. */
#define ORC_IDX(i, j, k, w, c) (((j) * (w) * (c)) + ((i) * (c)) + (k))

/* Wrapped neighbour indices exactly as written at src/fluid.cpp:159-162. */
#define ORC_WRAP_MINUS(i, n) ((((i) - 1) % (n) + (n)) % (n))
#define ORC_WRAP_PLUS(i, n) (((i) + 1) % (n))

/* src/fluid.cpp:19-21: (xl+xr+xt+xb+alpha*b)/beta, summed left to right. */
static inline R NAME(jacobi)(R xl, R xr, R xt, R xb, R alpha, R beta, R b) {
    return (xl + xr + xt + xb + alpha * b) / beta;
}

/* Semi-Lagrangian back-trace of channels 0,1 (src/fluid.cpp:24-70). */
void NAME(advect)(FIELD *vp, FIELD *vp_out, R dt) {
    const int width = vp->x, height = vp->y, depth = vp->z;
    const R *field = vp->data;
    R *new_field = vp_out->data;
    const R fwidth = (R)width;
    const R fheight = (R)height;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < height; j++) {
        for (int i = 0; i < width; i++) {
            const int x_idx = ORC_IDX(i, j, 0, width, depth);
            const int y_idx = ORC_IDX(i, j, 1, width, depth);
            R x_prev = (R)i - dt * field[x_idx] / fwidth;            /* :39 */
            R y_prev = (R)j - dt * field[y_idx] / fheight;           /* :41 */
            x_prev = FMOD(FMOD(x_prev, fwidth) + fwidth, fwidth);    /* :48 */
            y_prev = FMOD(FMOD(y_prev, fheight) + fheight, fheight); /* :49 */
            const int i0 = (int)x_prev, j0 = (int)y_prev;            /* :55 */
            const int i1 = (i0 + 1) % width, j1 = (j0 + 1) % height; /* :57 */
            const R sx = x_prev - (R)i0, sy = y_prev - (R)j0;        /* :58 */
            for (int k = 0; k < 2; k++) {                            /* :61-67 */
                new_field[ORC_IDX(i, j, k, width, depth)] =
                    (1 - sx) * (1 - sy) * field[ORC_IDX(i0, j0, k, width, depth)] +
                    sx * (1 - sy) * field[ORC_IDX(i1, j0, k, width, depth)] +
                    (1 - sx) * sy * field[ORC_IDX(i0, j1, k, width, depth)] +
                    sx * sy * field[ORC_IDX(i1, j1, k, width, depth)];
            }
        }
    }
}

/* Dye advection through the velocity field (src/fluid.cpp:72-127). */
void NAME(advect_color)(FIELD *image, FIELD *itmp, FIELD *vp, R dt) {
    const int iwidth = image->x, iheight = image->y, idepth = image->z;
    const int vwidth = vp->x, vdepth = vp->z;
    const R fiwidth = (R)iwidth, fiheight = (R)iheight;
    const R fvwidth = (R)vwidth, fvheight = (R)vp->y;
    const R viw = (fvwidth / fiwidth);   /* :82 */
    const R vih = (fvheight / fiheight); /* :83 */
    const R *img = image->data;
    const R *vel = vp->data;
    R *out = itmp->data;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < iheight; j++) {
        for (int i = 0; i < iwidth; i++) {
            const int vi = (int)((R)i * viw); /* :89 */
            const int vj = (int)((R)j * vih); /* :90 */
            const int vx_idx = ORC_IDX(vi, vj, 0, vwidth, vdepth);
            const int vy_idx = ORC_IDX(vi, vj, 1, vwidth, vdepth);
            R x_prev = (R)i - (dt / viw) * vel[vx_idx] / fiwidth;       /* :97 */
            R y_prev = (R)j - (dt / vih) * vel[vy_idx] / fiheight;      /* :98 */
            x_prev = FMOD(FMOD(x_prev, fiwidth) + fiwidth, fiwidth);     /* :105 */
            y_prev = FMOD(FMOD(y_prev, fiheight) + fiheight, fiheight);  /* :106 */
            const int i0 = (int)x_prev, j0 = (int)y_prev;                /* :112 */
            const int i1 = (i0 + 1) % iwidth, j1 = (j0 + 1) % iheight;  /* :113 */
            const R sx = x_prev - (R)i0, sy = y_prev - (R)j0;            /* :114 */
            for (int k = 0; k < idepth; k++) {                           /* :117-124 */
                out[ORC_IDX(i, j, k, iwidth, idepth)] =
                    (1 - sx) * (1 - sy) * img[ORC_IDX(i0, j0, k, iwidth, idepth)] +
                    sx * (1 - sy) * img[ORC_IDX(i1, j0, k, iwidth, idepth)] +
                    (1 - sx) * sy * img[ORC_IDX(i0, j1, k, iwidth, idepth)] +
                    sx * sy * img[ORC_IDX(i1, j1, k, iwidth, idepth)];
            }
        }
    }
}

/*
 * K Jacobi smoothing sweeps of channels 0,1 (src/fluid.cpp:129-196).  The
 * struct data pointers are swapped after every sweep but the last
 * (src/fluid.cpp:188-194), so the caller's structs change identity.
 */
void NAME(diffuse)(FIELD *vp, FIELD *vp_out, R viscosity, R dt, int iters) {
    const R alpha = viscosity * dt;                  /* :144 */
    const R beta = (R)(1.0 + 4.0 * (double)alpha);   /* :145, evaluated in double */
    const int w = vp->x, h = vp->y, c = vp->z;
    for (int iter = 0; iter < iters; iter++) {
        const R *data = vp->data;
        R *data_out = vp_out->data;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < h; j++) {
            for (int i = 0; i < w; i++) {
                const int iminus = ORC_WRAP_MINUS(i, w);
                const int iplus = ORC_WRAP_PLUS(i, w);
                const int jminus = ORC_WRAP_MINUS(j, h);
                const int jplus = ORC_WRAP_PLUS(j, h);
                for (int k = 0; k < 2; k++) {
                    const R left = data[ORC_IDX(iminus, j, k, w, c)];
                    const R right = data[ORC_IDX(iplus, j, k, w, c)];
                    const R top = data[ORC_IDX(i, jminus, k, w, c)];
                    const R bottom = data[ORC_IDX(i, jplus, k, w, c)];
                    const R u_n = data[ORC_IDX(i, j, k, w, c)];
                    data_out[ORC_IDX(i, j, k, w, c)] =
                        NAME(jacobi)(alpha * left, alpha * right, alpha * top,
                                     alpha * bottom, (R)1.0f, beta, u_n);
                }
            }
        }
        if (iter != iters - 1) { /* :188 */
            R *tp = vp_out->data;
            vp_out->data = vp->data;
            vp->data = tp;
        }
    }
}

/* Empty in the reference (src/fluid.cpp:198-208). */
void NAME(add_forces)(FIELD *vp, R *forces) {
    (void)vp;
    (void)forces;
}

/*
 * Force injection, filling in the reference's empty loop over every cell
 * and channel (src/fluid.cpp:198-208; "forces - A 3D array of force values
 * for each pixel in the grid", includes/fluid.hpp:61-66) with the explicit
 * Euler update u += dt * f of the first fz channels, forces laid out
 * [y][x][fz].  No reference behaviour to match: the library's
 * pfs_add_forces_apply is checked against this.
 */
void NAME(add_forces_apply)(FIELD *vp, const R *forces, int fz, R dt) {
    const int w = vp->x, h = vp->y, c = vp->z;
    for (int x = 0; x < w; x++)
        for (int y = 0; y < h; y++)
            for (int z = 0; z < fz; z++) {
                const R f = dt * forces[(y * w + x) * fz + z];
                vp->data[ORC_IDX(x, y, z, w, c)] = vp->data[ORC_IDX(x, y, z, w, c)] + f;
            }
}

/*
 * Divergence into channel 3 of both buffers, then K Jacobi sweeps of the
 * pressure Poisson equation on channel 2 (src/fluid.cpp:210-267).
 */
void NAME(compute_pressure)(FIELD *vp, FIELD *vp_out, R dt, int iters) {
    const int w = vp->x, h = vp->y, d = vp->z;
    const R alpha = (R)1.0f;                   /* :215 */
    const R beta = (R)4.0f;                    /* :216 */
    const R gamma = (R)(-1.0 / (double)dt);    /* :218 */
    {
        R *data_in = vp->data;
        R *data_o = vp_out->data;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < h; j++) {
            for (int i = 0; i < w; i++) {
                const int iminus = ORC_WRAP_MINUS(i, w);
                const int iplus = ORC_WRAP_PLUS(i, w);
                const int jminus = ORC_WRAP_MINUS(j, h);
                const int jplus = ORC_WRAP_PLUS(j, h);
                const R uR = data_in[ORC_IDX(iplus, j, 0, w, d)] - data_in[ORC_IDX(iminus, j, 0, w, d)];
                const R vT = data_in[ORC_IDX(i, jplus, 1, w, d)] - data_in[ORC_IDX(i, jminus, 1, w, d)];
                data_in[ORC_IDX(i, j, 3, w, d)] = gamma * (uR + vT); /* :234 */
                data_o[ORC_IDX(i, j, 3, w, d)] = gamma * (uR + vT);  /* :235 */
            }
        }
    }
    for (int iter = 0; iter < iters; iter++) {
        const R *data_in = vp->data;
        R *data_o = vp_out->data;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < h; j++) {
            for (int i = 0; i < w; i++) {
                const int iminus = ORC_WRAP_MINUS(i, w);
                const int iplus = ORC_WRAP_PLUS(i, w);
                const int jminus = ORC_WRAP_MINUS(j, h);
                const int jplus = ORC_WRAP_PLUS(j, h);
                const R pL = data_in[ORC_IDX(iminus, j, 2, w, d)];
                const R pR = data_in[ORC_IDX(iplus, j, 2, w, d)];
                const R pT = data_in[ORC_IDX(i, jminus, 2, w, d)];
                const R pB = data_in[ORC_IDX(i, jplus, 2, w, d)];
                const R b = data_in[ORC_IDX(i, j, 3, w, d)];
                data_o[ORC_IDX(i, j, 2, w, d)] = NAME(jacobi)(pL, pR, pT, pB, alpha, beta, b);
            }
        }
        if (iter != iters - 1) { /* :260 */
            R *tp = vp_out->data;
            vp_out->data = vp->data;
            vp->data = tp;
        }
    }
}

/* Projection: velocity minus the central-difference pressure gradient (src/fluid.cpp:269-296). */
void NAME(subtract_pressure_gradient)(FIELD *vp, FIELD *vp_out, R dt) {
    const int w = vp->x, h = vp->y, d = vp->z;
    const R *data_in = vp->data;
    R *data_o = vp_out->data;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < h; j++) {
        for (int i = 0; i < w; i++) {
            const int iminus = ORC_WRAP_MINUS(i, w);
            const int iplus = ORC_WRAP_PLUS(i, w);
            const int jminus = ORC_WRAP_MINUS(j, h);
            const int jplus = ORC_WRAP_PLUS(j, h);
            const R pL = data_in[ORC_IDX(iminus, j, 2, w, d)];
            const R pR = data_in[ORC_IDX(iplus, j, 2, w, d)];
            const R pT = data_in[ORC_IDX(i, jminus, 2, w, d)];
            const R pB = data_in[ORC_IDX(i, jplus, 2, w, d)];
            const R gradX = (pR - pL) * dt / (R)2.0f; /* :288 */
            const R gradY = (pB - pT) * dt / (R)2.0f; /* :289 */
            data_o[ORC_IDX(i, j, 0, w, d)] = data_in[ORC_IDX(i, j, 0, w, d)] - gradX;
            data_o[ORC_IDX(i, j, 1, w, d)] = data_in[ORC_IDX(i, j, 1, w, d)] - gradY;
        }
    }
}

/* src/fluid.cpp:298-310: advect -> diffuse -> computePressure -> subtractPressureGradient. */
void NAME(simulate_fluid_step)(FIELD *vp, FIELD *tmp, R dt, R viscosity, int iters) {
    NAME(advect)(vp, tmp, dt);
    NAME(diffuse)(tmp, vp, viscosity, dt, iters);
    NAME(compute_pressure)(vp, tmp, dt, iters);
    NAME(subtract_pressure_gradient)(tmp, vp, dt);
}

/* src/fluid.cpp:312-320: advect the dye, then swap image and itmp. */
void NAME(advect_color_step)(FIELD *image, FIELD *itmp, FIELD *vp, R dt) {
    NAME(advect_color)(image, itmp, vp, dt);
    R *tp = image->data;
    image->data = itmp->data;
    itmp->data = tp;
}

/*
 * The step loop of the reference driver (src/main.cpp:219-240, CPU branch):
 * simulate_fluid_step then advect_color_step, `steps` times.
 */
void NAME(run)(FIELD *vp, FIELD *tmp, FIELD *image, FIELD *itmp, R dt, R viscosity,
               int iters, int steps) {
    for (int s = 0; s < steps; s++) {
        NAME(simulate_fluid_step)(vp, tmp, dt, viscosity, iters);
        NAME(advect_color_step)(image, itmp, vp, dt);
    }
}

#undef ORC_IDX
#undef ORC_WRAP_MINUS
#undef ORC_WRAP_PLUS
