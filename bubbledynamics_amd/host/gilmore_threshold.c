/* Host program of the inertial-cavitation threshold table, written in C
 * against the C ABI the way the reference's Gilmore main() (G:433-631) drives
 * its kernel:
 *   - liver constants and the controller: bd_gilmore_reference_params (the
 *     __constant__ copies of G:440-489);
 *   - the 55 log-spaced tones of 20 kHz - 5 MHz and their triplets
 *     (G:465-469, G:498-520): 55 single + 1,485 dual + 26,235 triple;
 *   - the sweep launch (G:573) over every R0 of the table for the chosen
 *     triplets: bd_gilmore_threshold_sweep;
 *   - the table file in the reference's format (G:580-606):
 *     bd_gilmore_sweep_write.
 *
 *   bd_gilmore_threshold OUT [FIRST_COMBO N_COMBOS [N_R0 [ARITHMETIC [TOL]]]]
 *
 * FIRST_COMBO/N_COMBOS pick a contiguous part of the 27,775 triplets (all by
 * default), N_R0 the radii 0.1 um + i * 0.1 um (100 by default), ARITHMETIC
 * 0 the reference's expression order (the default), 1 the fast one, TOL the
 * DP5(4) absolute and relative tolerance (the reference's 1e-9 by default).  The
 * table written covers the chosen triplets.  Exit status: 0, or the library's
 * status code of the first call that failed. */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "bubbledynamics.h"

#define N_FR 55

static int hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return 1;
    fprintf(stderr, "bd_gilmore_threshold: %s: %s\n", what, hipGetErrorString(e));
    return 0;
}

/* Triplets (f_i, f_ii, f_k): i <= ii <= k with ii == i only for the first ii
 * of each i (G:498-520); returns their number, fills fr (3 per triplet) when
 * it is not NULL. */
static long triplets(const double f[N_FR], double* fr) {
    long n = 0;
    for (int i = 0; i < N_FR; ++i) {
        for (int ii = i; ii < N_FR; ++ii) {
            const int kk = ii == i ? ii : ii + 1;
            for (int k = kk; k < N_FR; ++k) {
                if (fr) { fr[3 * n] = f[i]; fr[3 * n + 1] = f[ii]; fr[3 * n + 2] = f[k]; }
                ++n;
            }
        }
    }
    return n;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s OUT [FIRST_COMBO N_COMBOS [N_R0 [ARITHMETIC [TOL]]]]\n", argv[0]);
        return 2;
    }
    const char* out = argv[1];
    /* tones, log-spaced (G:465-469) */
    const double min_fr = 0.02e6, max_fr = 5e6;
    const double step_log = (log10(max_fr) - log10(min_fr)) / (N_FR - 1.0);
    double f[N_FR];
    for (int i = 0; i < N_FR; ++i) f[i] = pow(10.0, log10(min_fr) + i * step_log);
    const long n_all = triplets(f, NULL);
    double* fr_all = (double*)malloc(sizeof(double) * 3 * (size_t)n_all);
    if (!fr_all) return BD_ERR_ALLOC;
    triplets(f, fr_all);

    const long first = argc > 2 ? atol(argv[2]) : 0;
    const long n_combo = argc > 3 ? atol(argv[3]) : n_all - first;
    const int n_r0 = argc > 4 ? atoi(argv[4]) : 100;
    const int arith = argc > 5 ? atoi(argv[5]) : BD_GM_ARITH_REFERENCE;
    const double tol = argc > 6 ? atof(argv[6]) : 0.0;
    if (first < 0 || n_combo <= 0 || first + n_combo > n_all || n_r0 <= 0 || tol < 0.0 ||
        (arith != BD_GM_ARITH_REFERENCE && arith != BD_GM_ARITH_FAST)) {
        fprintf(stderr, "bd_gilmore_threshold: bad range (%ld triplets in all)\n", n_all);
        free(fr_all);
        return 2;
    }

    bd_gilmore_params prm;
    bd_gilmore_reference_params(&prm);
    prm.arithmetic = arith;
    if (tol > 0.0) prm.abs_eps = prm.rel_eps = tol;

    const size_t n_cell = (size_t)n_combo * (size_t)n_r0;
    double *d_fr = NULL, *d_res = NULL;
    double* res = (double*)malloc(sizeof(double) * n_cell);
    int rc = res ? BD_OK : BD_ERR_ALLOC;
    if (!rc && !hip_ok(hipMalloc((void**)&d_fr, sizeof(double) * 3 * (size_t)n_combo), "hipMalloc")) rc = BD_ERR_HIP;
    if (!rc && !hip_ok(hipMalloc((void**)&d_res, sizeof(double) * n_cell), "hipMalloc")) rc = BD_ERR_HIP;
    if (!rc && !hip_ok(hipMemcpy(d_fr, fr_all + 3 * first, sizeof(double) * 3 * (size_t)n_combo,
                                 hipMemcpyHostToDevice), "hipMemcpy")) rc = BD_ERR_HIP;
    struct timespec w0, w1;
    timespec_get(&w0, TIME_UTC);
    if (!rc) rc = bd_gilmore_threshold_sweep(&prm, d_res, d_fr, (int)n_combo, n_r0, 0, NULL);
    if (!rc && !hip_ok(hipMemcpy(res, d_res, sizeof(double) * n_cell, hipMemcpyDeviceToHost), "hipMemcpy"))
        rc = BD_ERR_HIP;
    timespec_get(&w1, TIME_UTC);
    const double seconds = (double)(w1.tv_sec - w0.tv_sec) + 1e-9 * (double)(w1.tv_nsec - w0.tv_nsec);
    if (!rc) rc = bd_gilmore_sweep_write(out, fr_all + 3 * first, (int)n_combo, res, n_r0, prm.minR0, prm.stepR0,
                                         seconds);
    if (!rc) printf("bd_gilmore_threshold: %ld triplets x %d radii in %.3f s -> %s\n", n_combo, n_r0, seconds, out);
    else fprintf(stderr, "bd_gilmore_threshold: failed with status %d\n", rc);
    if (d_fr) (void)hipFree(d_fr);
    if (d_res) (void)hipFree(d_res);
    free(res);
    free(fr_all);
    return rc;
}
