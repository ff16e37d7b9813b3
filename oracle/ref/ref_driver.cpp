/*
 * extern "C" entry points around the reference's compiled src/fluid.cpp so
 * that Python (ctypes) can drive it.  Linked only into oracle/_ref/*.so,
 * which are test infrastructure: golden-vector generation and the optional
 * "reference" CPU baseline.  Each function forwards to the reference's own
 * function of the same name (includes/fluid.hpp:32,45,58,81,92,109,118).
 */
extern "C" {

int ref_jacobi_iters(void) { return NUM_JACOBI_ITERS; }

void ref_advect(vp_field *vp, vp_field *out, float dt) { advect(vp, out, dt); }

void ref_advect_color(vp_field *image, vp_field *out, vp_field *vp, float dt) {
    advect_color(image, out, vp, dt);
}

void ref_diffuse(vp_field *vp, vp_field *out, float visc, float dt) { diffuse(vp, out, visc, dt); }

void ref_compute_pressure(vp_field *vp, vp_field *out, float dt) { computePressure(vp, out, dt); }

void ref_subtract_pressure_gradient(vp_field *vp, vp_field *out, float dt) {
    subtractPressureGradient(vp, out, dt);
}

void ref_simulate_fluid_step(vp_field *vp, vp_field *tmp, float dt, float visc) {
    simulate_fluid_step(vp, tmp, dt, visc);
}

void ref_advect_color_step(vp_field *image, vp_field *itmp, vp_field *vp, float dt) {
    advect_color_step(image, itmp, vp, dt);
}

/* The CPU step loop of src/main.cpp:219-240 (timing mode, no frames). */
void ref_run(vp_field *vp, vp_field *tmp, vp_field *image, vp_field *itmp, float dt, float visc,
             int steps) {
    for (int i = 0; i < steps; i++) {
        simulate_fluid_step(vp, tmp, dt, visc);
        advect_color_step(image, itmp, vp, dt);
    }
}
}
