// FP64 serial restatement of the reference's heat/dose update and trajectory.
//
// TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU
// baseline; the shipped library never links or calls it.
//
// Reference: Westervelt_Temperature_SquareSpiral.cu (cited as T:line).
//   Temperature_square (FTCS + shifted Q + CEM43)   T:52-105
//   coefficient precompute                          T:212-224
//   initial T / TD                                   T:233-257
//   square-spiral trajectory                         T:141-173, T:323-409
// Pennes perfusion and the Kahan dose are extensions (perf = 0 and no
// compensation reproduce the reference).

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/bubbledynamics.h"

extern "C" {

// T:52-105 (use_p: read T_p, write T_n), plus perfusion in tissue and an
// optional Kahan compensation of the dose (TDc != NULL).
void orc_temperature_square(const bd_bioheat_params* p, double* T_n, double* T_p, const double* Q,
                            double* TD, double* TDc, int Nxy, int Nz, int use_p, int shift_x, int shift_y,
                            int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const long long Nxy_2 = (long long)Nxy * Nxy;
    double* dst = use_p ? T_n : T_p;
    const double* src = use_p ? T_p : T_n;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz < Nz - 1; ++iz) {
        const double Z = iz * p->dz;                         // T:79
        const bool water = Z < p->water_tissue;
        const double kxy = water ? p->kdt_dxy_water : p->kdt_dxy;
        const double kz = water ? p->kdt_dz_water : p->kdt_dz;
        const double dq = water ? p->dt_rho_ct_water : p->dt_rho_ct;
        for (int iy = 1; iy < Nxy - 1; ++iy)
            for (int ix = 1; ix < Nxy - 1; ++ix) {
                const long long idx = iz * Nxy_2 + ix + (long long)iy * Nxy;
                const int ix_ac = ix - shift_x, iy_ac = iy - shift_y;
                double q_ac = 0.0;                            // T:75-77
                if (ix_ac >= 0 && ix_ac < Nxy && iy_ac >= 0 && iy_ac < Nxy)
                    q_ac = Q[iz * Nxy_2 + ix_ac + (long long)iy_ac * Nxy];
                const double left = src[idx - 1], right = src[idx + 1];
                const double up = src[idx + Nxy], down = src[idx - Nxy];
                const double zp = src[idx + Nxy_2], zm = src[idx - Nxy_2];
                double Tn = src[idx] * (1.0 - 4.0 * kxy - 2.0 * kz) + kxy * (left + right + up + down) +
                            kz * (zp + zm) + dq * q_ac;       // T:85-96
                if (!water && p->perf != 0.0) Tn = Tn - p->perf * (src[idx] - p->T_art);
                dst[idx] = Tn;
                const double T_c = Tn - 273.15;
                const double add = (T_c < 43.0) ? std::pow(4.0, T_c - 43.0) * p->dt_T
                                                 : std::pow(2.0, T_c - 43.0) * p->dt_T;  // T:99-102
                if (!TDc) {
                    TD[idx] = TD[idx] + add;
                } else {
                    const double y = add - TDc[idx], t = TD[idx] + y;
                    TDc[idx] = (t - TD[idx]) - y;
                    TD[idx] = t;
                }
            }
    }
}

// T:141-173 and T:323-409: integer shifts per time step of the continuous spiral.
int orc_spiral_reference(double speed, double scale, double speed_init, double t_acc, double dt, double dxy,
                         double start, int spiral_lines, int* out_x, int* out_y, int cap) {
    const int Nt_scale = 1 + (int)(scale / (speed * dt));
    const int Nt_acc = 1 + (int)(t_acc / dt);
    const double acceleration = (speed - speed_init) / t_acc;
    std::vector<double> speed_acc(Nt_acc);
    for (int i = 0; i < Nt_acc; i++) speed_acc[i] = speed_init + i * dt * acceleration;
    double x_s = start, y_s = start;                         // T:331-332
    int sign_x = 1, sign_y = -1, shift_x = 0, shift_y = 0;
    bool right_left = true;
    double x_next = x_s, x_prev = x_s, y_next = y_s, y_prev = y_s;
    double speed_current = speed_init;
    int lines = 0, increase = 1, count = 0;
    while (lines < spiral_lines) {
        int n = 0;
        if (lines == 0) n = 1;
        x_next = x_prev + sign_x * dxy;
        y_next = y_prev + sign_y * dxy;
        for (; n < increase * Nt_scale + 2 * Nt_acc; n++) {
            if (n < Nt_acc) speed_current = speed_acc[n];
            if (n >= Nt_acc && n < increase * Nt_scale + Nt_acc) speed_current = speed;
            if (n >= increase * Nt_scale + Nt_acc) speed_current = speed_acc[increase * Nt_scale + 2 * Nt_acc - n - 1];
            if (right_left) {
                x_s = x_s + sign_x * speed_current * dt;
                const double diff_prev = std::fabs(x_s - x_prev) / dxy;
                const double diff_next = std::fabs(x_next - x_s) / dxy;
                if (diff_next < diff_prev) {
                    shift_x = shift_x + sign_x;
                    x_prev = x_next;
                    x_next = x_next + sign_x * dxy;
                }
            } else {
                y_s = y_s + sign_y * speed_current * dt;
                const double diff_prev = std::fabs(y_s - y_prev) / dxy;
                const double diff_next = std::fabs(y_next - y_s) / dxy;
                if (diff_next < diff_prev) {
                    shift_y = shift_y + sign_y;
                    y_prev = y_next;
                    y_next = y_next + sign_y * dxy;
                }
            }
            if (count < cap) { out_x[count] = shift_x; out_y[count] = shift_y; }
            ++count;
        }
        right_left = !right_left;
        if (right_left) {
            increase++;
            sign_x = -1 * sign_x;
            x_next = x_prev;
        } else {
            sign_y = -1 * sign_y;
            y_next = y_prev;
        }
        lines++;
    }
    return count;
}

}  // extern "C"
