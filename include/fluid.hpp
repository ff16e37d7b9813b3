/*
 * fluid.hpp — drop-in replacement for the reference's includes/fluid.hpp,
 * implemented by libpfs.so on an MI355X (gfx950) GPU.
 *
 * Same type, same function names, same argument meaning as the reference
 * header (includes/fluid.hpp:17-22 and :29-119), with results identical to
 * the reference's CPU solver src/fluid.cpp.  Differences from the reference
 * header:
 *   - No #ifdef USE_CUDA: both the vp_field flavour (:32-118) and the raw
 *     device-pointer flavour of the two step functions (:107, :116) are
 *     declared together; they are distinct C++ overloads.
 *   - The vp_field flavour accepts host memory (as the reference's CPU build
 *     does: data is staged to the GPU and back, and the call returns with the
 *     results in host memory) or device memory (work is enqueued on the null
 *     stream and the call returns at once, like the reference's CUDA build).
 *   - The __global__ per-operator kernels of the CUDA flavour (:30,43,56,68,
 *     79,90) are launched only inside the reference's own fluid.cu and are
 *     not declared; per-operator calls on device data use the vp_field
 *     flavour.
 *   - The Jacobi count is a run-time setting, initially NUM_JACOBI_ITERS (30,
 *     as in the reference), changed with set_num_jacobi_iters().
 * Errors: the functions return void like the reference's; a failure is
 * printed to stderr and left in pfs_last_error() (include/pfs.h).
 */

#ifndef FLUID_HPP_
#define FLUID_HPP_

#define NUM_JACOBI_ITERS (30)  // initial Jacobi count, as includes/fluid.hpp:11

/* includes/fluid.hpp:17-22 */
typedef struct {
    int x;
    int y;
    int z;
    float *data;  // [y][x][z], host or device memory
} vp_field;

/* includes/fluid.hpp:32 */
void advect(vp_field *vp, vp_field *vp_out, float dt);
/* includes/fluid.hpp:45 */
void advect_color(vp_field *image, vp_field *out, vp_field *vp, float dt);
/* includes/fluid.hpp:58 */
void diffuse(vp_field *vp, vp_field *vp_out, float viscosity, float dt);
/* includes/fluid.hpp:70 */
void addForces(vp_field *vp, float *forces);
/* includes/fluid.hpp:81 */
void computePressure(vp_field *vp, vp_field *vp_out, float dt);
/* includes/fluid.hpp:92 */
void subtractPressureGradient(vp_field *vp, vp_field *vp_out, float dt);
/* includes/fluid.hpp:109 */
void simulate_fluid_step(vp_field *vp, vp_field *tmp, float dt, float viscosity);
/* includes/fluid.hpp:118 */
void advect_color_step(vp_field *image, vp_field *itmp, vp_field *vp, float dt);

/* includes/fluid.hpp:107 — device pointers, result left in *vp. */
void simulate_fluid_step(float **vp, float **tmp, float dt, float viscosity, int vx, int vy,
                         int vz);
/* includes/fluid.hpp:116 — device pointers, swaps *image and *itmp. */
void advect_color_step(float **image, float **itmp, float **vp, float dt, int ix, int iy, int iz,
                       int vx, int vy, int vz);

/* Run-time counterpart of the reference's compile-time NUM_JACOBI_ITERS. */
void set_num_jacobi_iters(int iters);
int get_num_jacobi_iters(void);

#endif  // FLUID_HPP_
