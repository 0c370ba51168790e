/* Host program of the thermal dose along the square spiral, written in C
 * against the C ABI the way the reference's thermal main() (T:109-683) drives
 * its kernel, on a reduced grid:
 *   - the reference's water and tissue constants and coefficients
 *     (T:115-173, T:194-224) on an NXY x NXY x NZ grid of its 0.2 / 0.1 mm
 *     cells centred on the axis;
 *   - the acoustic heat source handed over as the reference does: a quarter-
 *     domain Q written as a Q3D file (the Westervelt program's writer,
 *     W:778-790) and read back with the quarter-to-full mirror expansion
 *     (T:269-306); here Q is a Gaussian focus on the axis;
 *   - the initial temperature and dose by region (T:233-257);
 *   - the continuous spiral's shift sequence (T:141-173, T:323-409) and, per
 *     time step, the drop-in Temperature_square launch with the ping-pong of
 *     T:382-390 (use_p alternating);
 *   - the end-of-run T and TD slices in the reference's text format
 *     (T:412-447) when an output directory is given;
 *   - then the same substeps through bd_bioheat_run (runs of equal shift,
 *     two substeps per launch) from the same initial state, and the largest
 *     differences of T and of the dose.
 *
 *   bd_temperature_spiral NXY NZ SUBSTEPS [OUTDIR]
 *
 * Prints "max|T| ... max|dT| ... max|dTD|/max|TD| ..." and exits 0, or with
 * the library's status code of the first call that failed. */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bubbledynamics.h"

#define CHECK_BD(call)                                                               \
    do {                                                                             \
        const int rc_ = (call);                                                      \
        if (rc_) { fprintf(stderr, "bd_temperature_spiral: %s -> %d\n", #call, rc_); \
                   return rc_; }                                                     \
    } while (0)
#define CHECK_HIP(call)                                                              \
    do {                                                                             \
        const hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) { fprintf(stderr, "bd_temperature_spiral: %s: %s\n",   \
                                        #call, hipGetErrorString(e_));               \
                                return BD_ERR_HIP; }                                 \
    } while (0)

/* The reference's thermal constants (T:115-173, T:212-224). */
static void params(bd_bioheat_params* p) {
    memset(p, 0, sizeof(*p));
    const double rho_w = 1000.0, c_w = 4200.0, K_w = 0.6;     /* T:125-128 */
    const double rho_t = 1055.0, c_t = 3700.0, K_t = 0.51;    /* T:132-135 */
    const double k_w = K_w / (rho_w * c_w), k_t = K_t / (rho_t * c_t);
    p->water_tissue = 0.12;                                   /* T:139 */
    p->dxy = 0.0002;                                          /* T:157 */
    p->dz = 0.0001;                                           /* T:162 */
    p->dt_T = 0.001;                                          /* T:166 */
    p->kdt_dxy = p->dt_T * k_t / (p->dxy * p->dxy);          /* T:212-217 */
    p->kdt_dz = p->dt_T * k_t / (p->dz * p->dz);
    p->kdt_dxy_water = p->dt_T * k_w / (p->dxy * p->dxy);
    p->kdt_dz_water = p->dt_T * k_w / (p->dz * p->dz);
    p->dt_rho_ct = p->dt_T / (rho_t * c_t);
    p->dt_rho_ct_water = p->dt_T / (rho_w * c_w);
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s NXY NZ SUBSTEPS [OUTDIR]\n", argv[0]);
        return 2;
    }
    const int nxy = atoi(argv[1]), nz = atoi(argv[2]), substeps = atoi(argv[3]);
    const char* outdir = argc > 4 ? argv[4] : NULL;
    if (nxy < 11 || nxy % 2 == 0 || nz < 10 || substeps < 1) {
        fprintf(stderr, "bd_temperature_spiral: NXY odd >= 11, NZ >= 10, SUBSTEPS >= 1\n");
        return 2;
    }
    bd_bioheat_params prm;
    params(&prm);
    const size_t n = (size_t)nxy * nxy * nz, bytes = n * sizeof(double);

    /* heat source: a quarter domain (axis at ix = iy = 0) holding a Gaussian
     * focus 60% down the grid, handed over through a Q3D file whose planes
     * start at iz0 = 2 like the Westervelt program's */
    const int nq = nxy / 2 + 1, iz0 = 2, nzq = nz + iz0;
    double* hq = (double*)calloc((size_t)nq * nq * nzq, sizeof(double));
    if (!hq) return BD_ERR_ALLOC;
    const double zf = 0.6 * nz + iz0, s2 = 2.0 * 3.0 * 3.0;
    for (int iz = 0; iz < nzq; ++iz)
        for (int iy = 0; iy < nq; ++iy)
            for (int ix = 0; ix < nq; ++ix) {
                const double r2 = (double)(ix * ix + iy * iy) + 0.25 * (iz - zf) * (iz - zf);
                hq[((size_t)iz * nq + iy) * nq + ix] = 1.0e8 * exp(-r2 / s2);
            }
    void *dq = NULL, *Q = NULL;
    CHECK_HIP(hipMalloc(&dq, (size_t)nq * nq * nzq * sizeof(double)));
    CHECK_HIP(hipMemcpy(dq, hq, (size_t)nq * nq * nzq * sizeof(double), hipMemcpyHostToDevice));
    free(hq);
    char q3d[4096];
    snprintf(q3d, sizeof q3d, "%s/Q3D_spiral.bin", outdir ? outdir : ".");
    CHECK_BD(bd_q3d_write(q3d, BD_F64, dq, nq, nzq, iz0, NULL));            /* W:778-790 */
    CHECK_HIP(hipMalloc(&Q, bytes));
    CHECK_BD(bd_q3d_read_expand(q3d, BD_F64, Q, nxy, nz, NULL));            /* T:269-306 */
    (void)hipFree(dq);
    if (!outdir) remove(q3d);

    /* initial state by region (T:233-257): T0 in both buffers, TD one step of dose */
    const double T0_t = 23.6, T0_w = 22.7;                                   /* T:149-150 */
    int iz_t = 0;
    while (iz_t < nz && iz_t * prm.dz < prm.water_tissue) ++iz_t;
    double* h = (double*)malloc(bytes);
    double* hd = (double*)malloc(bytes);
    if (!h || !hd) return BD_ERR_ALLOC;
    for (size_t i = 0; i < n; ++i) {
        const int water = (int)(i / ((size_t)nxy * nxy)) < iz_t;
        const double T0 = water ? T0_w : T0_t;
        h[i] = T0 + 273.15;
        hd[i] = pow(4.0, T0 - 43.0) * prm.dt_T;
    }
    void *Ta = NULL, *Tb = NULL, *TD = NULL, *Ua = NULL, *Ub = NULL, *UD = NULL;
    CHECK_HIP(hipMalloc(&Ta, bytes)); CHECK_HIP(hipMalloc(&Tb, bytes)); CHECK_HIP(hipMalloc(&TD, bytes));
    CHECK_HIP(hipMalloc(&Ua, bytes)); CHECK_HIP(hipMalloc(&Ub, bytes)); CHECK_HIP(hipMalloc(&UD, bytes));
    void* init[6] = {Ta, Tb, Ua, Ub, TD, UD};
    for (int i = 0; i < 6; ++i) CHECK_HIP(hipMemcpy(init[i], i < 4 ? h : hd, bytes, hipMemcpyHostToDevice));

    /* the spiral's shifts (T:141-173, T:323-409) */
    const double speed = 2e-3, scale = 3e-3, speed_init = 0.05e-3, t_acc = 0.15, start = -0.085 + 425 * 2e-4;
    const int total = bd_spiral_reference(speed, scale, speed_init, t_acc, prm.dt_T, prm.dxy, start, 13, NULL,
                                          NULL, 0);
    if (total < substeps) return BD_ERR_INVALID;
    int* sx = (int*)malloc(sizeof(int) * (size_t)total);
    int* sy = (int*)malloc(sizeof(int) * (size_t)total);
    if (!sx || !sy) return BD_ERR_ALLOC;
    bd_spiral_reference(speed, scale, speed_init, t_acc, prm.dt_T, prm.dxy, start, 13, sx, sy, total);

    /* the reference's loop: one Temperature_square launch per time step (T:382-390) */
    int use_p = 1;                                  /* T_p holds the current temperature */
    for (int s = 0; s < substeps; ++s) {
        CHECK_BD(bd_temperature_square(&prm, BD_F64, Tb, Ta, Q, TD, nxy, nz, use_p, sx[s], sy[s], NULL));
        use_p = !use_p;
    }
    void* Tcur = use_p ? Ta : Tb;

    /* the library's loop: runs of equal shift, two substeps per launch */
    int which = 0;
    for (int s = 0; s < substeps;) {
        int e = s + 1;
        while (e < substeps && sx[e] == sx[s] && sy[e] == sy[s]) ++e;
        CHECK_BD(bd_bioheat_run(&prm, BD_F64, Ua, Ub, Q, UD, NULL, nxy, nz, sx[s], sy[s], e - s, &which, NULL));
        s = e;
    }
    void* Ucur = which ? Ub : Ua;
    CHECK_HIP(hipDeviceSynchronize());

    if (outdir) {                                   /* T:412-447: the end-of-run slices */
        char path[4096];
        snprintf(path, sizeof path, "%s/T_xz.txt", outdir);
        CHECK_BD(bd_write_slice(path, BD_F64, Tcur, nxy, nz, BD_SLICE_XZ, nxy / 2, 0, nz, 0, nxy, 0.0, NULL));
        snprintf(path, sizeof path, "%s/TD_xz.txt", outdir);
        CHECK_BD(bd_write_slice(path, BD_F64, TD, nxy, nz, BD_SLICE_XZ, nxy / 2, 0, nz, 0, nxy, 0.0, NULL));
    }

    double* u = (double*)malloc(bytes);
    double* ud = (double*)malloc(bytes);
    if (!u || !ud) return BD_ERR_ALLOC;
    CHECK_HIP(hipMemcpy(h, Tcur, bytes, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(u, Ucur, bytes, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hd, TD, bytes, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(ud, UD, bytes, hipMemcpyDeviceToHost));
    double tmax = 0.0, dt = 0.0, dmax = 0.0, ddose = 0.0;
    int finite = 1;
    for (size_t i = 0; i < n; ++i) {
        if (!isfinite(h[i]) || !isfinite(u[i]) || !isfinite(hd[i]) || !isfinite(ud[i])) finite = 0;
        if (h[i] > tmax) tmax = h[i];
        if (fabs(h[i] - u[i]) > dt) dt = fabs(h[i] - u[i]);
        if (hd[i] > dmax) dmax = hd[i];
        if (fabs(hd[i] - ud[i]) > ddose) ddose = fabs(hd[i] - ud[i]);
    }
    printf("substeps %d of %d max|T| %.17g max|dT| %.17g max|TD| %.17g max|dTD| %.17g finite %d\n", substeps, total,
           tmax, dt, dmax, ddose, finite);
    void* all[7] = {Q, Ta, Tb, TD, Ua, Ub, UD};
    for (int i = 0; i < 7; ++i) (void)hipFree(all[i]);
    free(h); free(hd); free(u); free(ud); free(sx); free(sy);
    return finite ? BD_OK : BD_ERR_NONFINITE;
}
