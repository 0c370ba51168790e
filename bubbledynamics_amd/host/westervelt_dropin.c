/* Host program of the Westervelt step loop, written in C against the C ABI
 * the way the reference's main() (W:350-843) runs it, on a small grid:
 *   - the reference's media, relaxation, transducer and Mur boundaries
 *     (W:356-417) on a coarsened grid that holds source, interface and focus
 *     (cap shrunk to the grid, interface at the middle plane);
 *   - per step the interior (W:560), t += dt (W:563), the hard source
 *     (W:566), the Mur faces (W:569) and the rotation of the six pressure and
 *     2 x 3 relaxation levels by nine device copies (W:575-584), each a
 *     drop-in call of the library;
 *   - then the same steps through the plan (pointer ring, one fused kernel
 *     per step, 6-step hipGraphs), and the largest difference of p^n.
 *
 *   bd_westervelt_dropin NXY NZ STEPS [f32|f64]
 *
 * Prints "max|p| ... max|dp| ..." and exits 0, or with the library's status
 * code of the first call that failed. */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bubbledynamics.h"

#define CHECK_BD(call)                                                              \
    do {                                                                            \
        const int rc_ = (call);                                                     \
        if (rc_) { fprintf(stderr, "bd_westervelt_dropin: %s -> %d\n", #call, rc_); \
                   return rc_; }                                                    \
    } while (0)
#define CHECK_HIP(call)                                                             \
    do {                                                                            \
        const hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) { fprintf(stderr, "bd_westervelt_dropin: %s: %s\n",   \
                                        #call, hipGetErrorString(e_));              \
                                return BD_ERR_HIP; }                                \
    } while (0)

static const double kPi = 3.14159265358979323846;

/* The reference's constants (W:356-417) on an nz-plane grid of 0.15 mm cells. */
static void params(bd_westervelt_params* p, int nz) {
    memset(p, 0, sizeof(*p));
    const double fr_ref = 1.253e6;                            /* W:356 */
    p->rho = 1000.0; p->c_inf = 1500.0;                       /* W:366-367 */
    p->abs_water = 0.026 * (fr_ref / 1e6) * (fr_ref / 1e6);   /* W:371 */
    p->delta = 2.0 * p->abs_water * p->c_inf * p->c_inf * p->c_inf /
               ((2.0 * kPi * fr_ref) * (2.0 * kPi * fr_ref));   /* W:372 */
    p->beta = 3.5;                                            /* W:373 */
    p->rho_tissue = 1055.0; p->c_inf_tissue = 1600.0;         /* W:376-377 */
    p->beta_tissue = 4.5;                                     /* W:381 */
    p->abs_tissue = 7.2 * fr_ref / 1e6;                       /* W:382 */
    p->delta_tissue = 4.2318e-5;                              /* W:389 */
    p->c1 = 4.03536; p->c2 = 6.14150;                         /* W:390-391 */
    p->tau1 = 1.6178e-8; p->tau2 = 1.1494e-7;                 /* W:392-393 */
    /* coarsened grid: 1 MHz, 0.15 mm cells, source plane at z = 0 */
    p->fr = 1.0e6;
    p->dxy = p->dz = 1.5e-4;
    p->z_min = -2.0 * p->dz;
    p->xy_min = 0.0;
    p->dt = 0.3 * p->dz / p->c_inf_tissue;
    p->water_tissue = p->z_min + (nz / 2) * p->dz - 0.5 * p->dz;
    p->d = 0.6 * (nz - 3) * p->dz;
    p->alpha = asin(0.45);
    p->alpha_hole = asin(0.1);
    p->p0 = 1.0e6;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s NXY NZ STEPS [f32|f64]\n", argv[0]);
        return 2;
    }
    const int nxy = atoi(argv[1]), nz = atoi(argv[2]), steps = atoi(argv[3]);
    const int dtype = (argc > 4 && strcmp(argv[4], "f32") == 0) ? BD_F32 : BD_F64;
    const size_t es = dtype == BD_F64 ? 8 : 4;
    if (nxy < 10 || nz < 10 || steps < 1) return 2;
    bd_westervelt_params prm;
    params(&prm, nz);
    const size_t n = (size_t)nxy * nxy * nz, bytes = n * es;

    /* the reference's twelve fields (W:523-534), zero at t = 0 */
    void* f[12];
    for (int i = 0; i < 12; ++i) {
        CHECK_HIP(hipMalloc(&f[i], bytes));
        CHECK_HIP(hipMemset(f[i], 0, bytes));
    }
    void *Pn = f[0], *P = f[1], *Pm1 = f[2], *Pm2 = f[3], *Pm3 = f[4], *Pm4 = f[5];
    void *R1 = f[6], *R1m1 = f[7], *R1m2 = f[8], *R2 = f[9], *R2m1 = f[10], *R2m2 = f[11];
    double t = 0.0;
    for (int s = 0; s < steps; ++s) {
        CHECK_BD(bd_westervelt_space6(&prm, dtype, Pn, P, Pm1, Pm2, Pm3, Pm4, R1, R1m1, R1m2, R2, R2m1, R2m2,
                                      nxy, nz, NULL));                                   /* W:560 */
        t += prm.dt;                                                                     /* W:563 */
        CHECK_BD(bd_boundary_condition_source(&prm, dtype, Pn, t, nxy, nz, NULL, NULL, NULL));  /* W:566 */
        CHECK_BD(bd_boundary_condition_space6(&prm, dtype, Pn, P, t, nxy, nz, NULL));    /* W:569 */
        /* W:575-584 */
        CHECK_HIP(hipMemcpy(Pm4, Pm3, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(Pm3, Pm2, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(Pm2, Pm1, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(Pm1, P, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(P, Pn, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(R1m2, R1m1, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(R1m1, R1, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(R2m2, R2m1, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipMemcpy(R2m1, R2, bytes, hipMemcpyDeviceToDevice));
    }

    /* the same steps through the plan */
    void* state = NULL;
    CHECK_HIP(hipMalloc(&state, bd_westervelt_state_elems(nxy, nz) * es));
    CHECK_HIP(hipMemset(state, 0, bd_westervelt_state_elems(nxy, nz) * es));
    bd_westervelt_plan* plan = NULL;
    CHECK_BD(bd_westervelt_plan_create(&prm, dtype, nxy, nz, steps, -1, -1, -1, &plan));
    CHECK_BD(bd_westervelt_plan_bind(plan, state, NULL, NULL, NULL, NULL, NULL));
    CHECK_BD(bd_westervelt_plan_run(plan, steps, 1, NULL));
    CHECK_HIP(hipDeviceSynchronize());
    const int slot = bd_westervelt_plan_slot(plan, 0, 0);

    double* a = (double*)malloc(n * sizeof(double));
    double* b = (double*)malloc(n * sizeof(double));
    void* h = malloc(bytes);
    if (!a || !b || !h) return BD_ERR_ALLOC;
    CHECK_HIP(hipMemcpy(h, P, bytes, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) a[i] = dtype == BD_F64 ? ((double*)h)[i] : (double)((float*)h)[i];
    CHECK_HIP(hipMemcpy(h, (char*)state + (size_t)slot * bytes, bytes, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) b[i] = dtype == BD_F64 ? ((double*)h)[i] : (double)((float*)h)[i];
    double amax = 0.0, dmax = 0.0;
    int finite = 1;
    for (size_t i = 0; i < n; ++i) {
        if (!isfinite(a[i]) || !isfinite(b[i])) finite = 0;
        if (fabs(a[i]) > amax) amax = fabs(a[i]);
        if (fabs(a[i] - b[i]) > dmax) dmax = fabs(a[i] - b[i]);
    }
    printf("steps %d t %.17g max|p| %.17g max|dp| %.17g finite %d\n", steps, t, amax, dmax, finite);
    CHECK_BD(bd_westervelt_plan_destroy(plan));
    for (int i = 0; i < 12; ++i) (void)hipFree(f[i]);
    (void)hipFree(state);
    free(a); free(b); free(h);
    return finite ? BD_OK : BD_ERR_NONFINITE;
}
