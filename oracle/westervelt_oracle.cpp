// FP64 serial restatement of the reference's Westervelt update formulae.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
// timed CPU baseline; the shipped library never links or calls it.
//
// Reference: Westervelt_relax_fft_Pressure_Energy.cu (cited as W:line).
//   interior update      W:139-236  -> orc_wv_space6
//   hard source plane    W:62-85    -> orc_wv_source
//   Mur boundaries       W:88-135   -> orc_wv_abc (defined outward order, see below)
//   water heat source    W:242-265  -> orc_acoustic_power_water
//   tissue harmonics     W:268-345  -> orc_harmonics_accumulate / orc_harmonics_q
//   step driver          W:556-585  -> orc_wv_run (pointer ring instead of 9 copies)
//
// The reference's boundary kernel reads neighbour cells that other threads of
// the same launch overwrite (W:104, W:112, W:125, W:132).  Here the faces are
// evaluated in a fixed order: all z faces (innermost layer first), then the x
// faces, then the y faces, every layer reading its inner neighbour's new value.
// That order also reproduces the reference's per-thread precedence at edges and
// corners (z, then x, then y; the last writer wins).
//
// Arithmetic follows the reference expression order so that results can be
// compared at rounding level; compiled with -ffp-contract=off, no fast-math.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/bubbledynamics.h"

namespace {

inline size_t IDX(int ix, int iy, int iz, int Nxy) {
    return (size_t)iz * Nxy * Nxy + (size_t)iy * Nxy + ix;
}

// First plane index whose z lies in tissue, with the reference's test
// z_min + iz*dz < water_tissue (W:212).
int first_tissue_plane(const bd_westervelt_params* p, int Nz) {
    for (int iz = 0; iz < Nz; ++iz)
        if (!(p->z_min + iz * p->dz < p->water_tissue)) return iz;
    return Nz;
}

// Sponge factor for an interior cell; 1.0 outside the damping layer.
double sponge(const bd_westervelt_params* p, int ix, int iy, int iz, int Nxy, int Nz) {
    const int W = p->sponge_cells;
    if (W <= 0) return 1.0;
    int dx = Nxy - 1 - ix, dy = Nxy - 1 - iy, dz = Nz - 1 - iz;
    if (p->full_domain) { dx = std::min(dx, ix); dy = std::min(dy, iy); }
    double s = 0.0;
    if (dx < W) { double a = p->sponge_alpha * (W - dx); s += a * a; }
    if (dy < W) { double a = p->sponge_alpha * (W - dy); s += a * a; }
    if (dz < W) { double a = p->sponge_alpha * (W - dz); s += a * a; }
    if (s == 0.0) return 1.0;
    return std::exp(-s);
}

// Value the source writes at column (ix, iy) of plane iz=2 at time t.
double source_value(const bd_westervelt_params* p, int ix, int iy, double t,
                    const double* amp, const double* phase) {
    const double X = p->xy_min + ix * p->dxy;
    const double Y = p->xy_min + iy * p->dxy;
    if (p->source == BD_SRC_ARRAY) {
        const double fx = std::floor((X - p->arr_x0) / p->arr_pitch);
        const double fy = std::floor((Y - p->arr_y0) / p->arr_pitch);
        if (fx < 0 || fy < 0 || fx >= p->arr_nx || fy >= p->arr_ny) return 0.0;
        const int e = (int)fy * p->arr_nx + (int)fx;
        double tb = t;
        if (p->burst_cycles > 0) {
            tb = t - p->burst_t0;
            if (tb < 0.0 || tb * p->fr >= p->burst_cycles) return 0.0;
        }
        return amp[e] * std::sin(2 * M_PI * p->fr * tb + phase[e]);
    }
    // W:72-83
    const double r = std::sqrt(X * X + Y * Y);
    if (r <= p->d * std::tan(p->alpha) && r >= p->d * std::tan(p->alpha_hole)) {
        const double var = std::sqrt(1.0 + std::pow(r / p->d, 2));
        double tt = t + (p->d * (var - 1.0) / p->c_inf);
        if (p->burst_cycles > 0) {
            tt = tt - p->burst_t0;
            if (tt < 0.0 || tt * p->fr >= p->burst_cycles) return 0.0;
        }
        return p->p0 * std::sin(2 * M_PI * p->fr * tt) / var;
    }
    return 0.0;
}

// Mur update, expression order of W:104.
inline double mur(double pm1_nb, double p_nb, double pm1_self, double k) {
    return pm1_nb + ((p_nb - pm1_self) * (k - 1.0) / (k + 1.0));
}

}  // namespace

extern "C" {

int orc_first_tissue_plane(const bd_westervelt_params* p, int Nz) { return first_tissue_plane(p, Nz); }

// W:139-236.  Energy partials (optional) are accumulated per plane into
// e_plane[2*iz], e_plane[2*iz+1].
static void wv_space6_impl(const bd_westervelt_params* p, double* Pnext, const double* P,
                           const double* Pm1, const double* Pm2, const double* Pm3,
                           const double* Pm4, double* R1, const double* R1m1,
                           const double* R1m2, double* R2, const double* R2m1,
                           const double* R2m2, int Nxy, int Nz, double* pmax,
                           double* e_plane) {
    const int itiss = first_tissue_plane(p, Nz);
    const int xlo = p->full_domain ? 3 : 0;
    const double dt = p->dt;
    const double dV = p->dxy * p->dxy * p->dz;
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz <= Nz - 4; ++iz) {
        double e0 = 0.0, e1 = 0.0;
        const bool water = iz < itiss;
        double c2dt2_dxy2, c2dt2_dz2, delta_2dtc2, beta2_rhoc2, rho_c, c_c;
        if (water) {  // W:214-217
            c2dt2_dxy2 = std::pow(p->c_inf * dt / p->dxy, 2);
            c2dt2_dz2 = std::pow(p->c_inf * dt / p->dz, 2);
            delta_2dtc2 = p->delta / (2.0 * dt * std::pow(p->c_inf, 2));
            beta2_rhoc2 = 2.0 * p->beta / (p->rho * std::pow(p->c_inf, 2));
            rho_c = p->rho; c_c = p->c_inf;
        } else {      // W:221-224
            c2dt2_dxy2 = std::pow(p->c_inf_tissue * dt / p->dxy, 2);
            c2dt2_dz2 = std::pow(p->c_inf_tissue * dt / p->dz, 2);
            delta_2dtc2 = p->delta_tissue / (2.0 * dt * std::pow(p->c_inf_tissue, 2));
            beta2_rhoc2 = 2.0 * p->beta_tissue / (p->rho_tissue * std::pow(p->c_inf_tissue, 2));
            rho_c = p->rho_tissue; c_c = p->c_inf_tissue;
        }
        for (int iy = xlo; iy <= Nxy - 4; ++iy) {
            for (int ix = xlo; ix <= Nxy - 4; ++ix) {
                const size_t idx = IDX(ix, iy, iz, Nxy);
                // even mirror about ix=0 / iy=0 (W:176-200): index -k reads +k
                auto mx = [&](int k) { int j = ix - k; if (j < 0) j = -j; return IDX(j, iy, iz, Nxy); };
                auto my = [&](int k) { int j = iy - k; if (j < 0) j = -j; return IDX(ix, j, iz, Nxy); };
                const double pc = P[idx];
                const double Dxx = (-490.0 * pc + 270.0 * (P[IDX(ix + 1, iy, iz, Nxy)] + P[mx(1)]) -
                                    27.0 * (P[IDX(ix + 2, iy, iz, Nxy)] + P[mx(2)]) +
                                    2.0 * (P[IDX(ix + 3, iy, iz, Nxy)] + P[mx(3)])) / 180.0;
                const double Dyy = (-490.0 * pc + 270.0 * (P[IDX(ix, iy + 1, iz, Nxy)] + P[my(1)]) -
                                    27.0 * (P[IDX(ix, iy + 2, iz, Nxy)] + P[my(2)]) +
                                    2.0 * (P[IDX(ix, iy + 3, iz, Nxy)] + P[my(3)])) / 180.0;
                const double Dzz = (-490.0 * pc + 270.0 * (P[IDX(ix, iy, iz + 1, Nxy)] + P[IDX(ix, iy, iz - 1, Nxy)]) -
                                    27.0 * (P[IDX(ix, iy, iz + 2, Nxy)] + P[IDX(ix, iy, iz - 2, Nxy)]) +
                                    2.0 * (P[IDX(ix, iy, iz + 3, Nxy)] + P[IDX(ix, iy, iz - 3, Nxy)])) / 180.0;
                // W:207-209
                const double Dttt = 5.0 * pc - 18.0 * Pm1[idx] + 24.0 * Pm2[idx] - 14.0 * Pm3[idx] + 3.0 * Pm4[idx];
                const double D2t2_p1 = std::pow(3.0 * pc - 4.0 * Pm1[idx] + Pm2[idx], 2) / 4.0;
                const double D2t2_p2 = pc * (2.0 * pc - 5.0 * Pm1[idx] + 4.0 * Pm2[idx] - Pm3[idx]);
                double sum_R = 0.0;
                if (!water) {  // W:226-230
                    const double Dttt_2dt = Dttt / (2.0 * std::pow(dt, 3));
                    R1[idx] = (2.0 * dt * 2.0 * p->c1 * p->tau1 * Dttt_2dt /
                               ((2.0 * dt + 3.0 * p->tau1) * std::pow(p->c_inf_tissue, 3))) +
                              ((p->tau1 * (4.0 * R1m1[idx] - R1m2[idx])) / (2.0 * dt + 3.0 * p->tau1));
                    R2[idx] = (2.0 * dt * 2.0 * p->c2 * p->tau2 * Dttt_2dt /
                               ((2.0 * dt + 3.0 * p->tau2) * std::pow(p->c_inf_tissue, 3))) +
                              ((p->tau2 * (4.0 * R2m1[idx] - R2m2[idx])) / (2.0 * dt + 3.0 * p->tau2));
                    sum_R = (R1[idx] + R2[idx]) * std::pow(p->c_inf_tissue * dt, 2);
                }
                // W:234
                double pn = 2.0 * pc - Pm1[idx] + c2dt2_dxy2 * (Dxx + Dyy) + c2dt2_dz2 * Dzz +
                            delta_2dtc2 * Dttt + beta2_rhoc2 * (D2t2_p1 + D2t2_p2) + sum_R;
                const double g = sponge(p, ix, iy, iz, Nxy, Nz);
                if (g != 1.0) pn = g * pn;
                Pnext[idx] = pn;
                if (pmax && pn > pmax[idx]) pmax[idx] = pn;
                if (e_plane) {
                    const double lap = c2dt2_dxy2 * (Dxx + Dyy) + c2dt2_dz2 * Dzz;
                    const double dp = pn - pc;
                    e0 += pn * pn * dV / (rho_c * c_c * c_c);
                    e1 += dp * dp - pn * lap;
                }
            }
        }
        if (e_plane) { e_plane[2 * iz] = e0; e_plane[2 * iz + 1] = e1; }
    }
}

void orc_wv_space6(const bd_westervelt_params* p, double* Pnext, const double* P,
                   const double* Pm1, const double* Pm2, const double* Pm3, const double* Pm4,
                   double* R1, const double* R1m1, const double* R1m2, double* R2,
                   const double* R2m1, const double* R2m2, int Nxy, int Nz) {
    wv_space6_impl(p, Pnext, P, Pm1, Pm2, Pm3, Pm4, R1, R1m1, R1m2, R2, R2m1, R2m2, Nxy, Nz,
                   nullptr, nullptr);
}

// W:62-85 (plane iz=2).
void orc_wv_source(const bd_westervelt_params* p, double* P, double t, int Nxy, int Nz,
                   const double* amp, const double* phase) {
    (void)Nz;
    for (int iy = 0; iy < Nxy; ++iy)
        for (int ix = 0; ix < Nxy; ++ix)
            P[IDX(ix, iy, 2, Nxy)] = source_value(p, ix, iy, t, amp, phase);
}

// W:88-135 in the defined outward order (see file header).  BD_BC_ZERO sets
// every boundary cell to zero except the source plane's interior columns.
void orc_wv_abc(const bd_westervelt_params* p, double* P, const double* Pm1, double t,
                int Nxy, int Nz) {
    (void)t;
    const int itiss = first_tissue_plane(p, Nz);
    const double dt = p->dt;
    const int lo = p->full_domain ? 3 : 0;
    if (p->boundary == BD_BC_ZERO) {
        for (int iz = 0; iz < Nz; ++iz)
            for (int iy = 0; iy < Nxy; ++iy)
                for (int ix = 0; ix < Nxy; ++ix) {
                    const bool inner_col = ix >= lo && ix <= Nxy - 4 && iy >= lo && iy <= Nxy - 4;
                    const bool interior = inner_col && iz >= 3 && iz <= Nz - 4;
                    if (interior || (inner_col && iz == 2)) continue;
                    P[IDX(ix, iy, iz, Nxy)] = 0.0;
                }
        return;
    }
    const double cdt_dz = p->c_inf * dt / p->dz;                   // W:100
    const double cdt_dz_t = p->c_inf_tissue * dt / p->dz;          // W:111
    // z faces, all columns
    for (int iy = 0; iy < Nxy; ++iy)
        for (int ix = 0; ix < Nxy; ++ix) {
            for (int iz = 1; iz >= 0; --iz) {     // z-min: iz=1 reads iz=2, then iz=0 reads iz=1
                const size_t i = IDX(ix, iy, iz, Nxy), zp = IDX(ix, iy, iz + 1, Nxy);
                P[i] = mur(Pm1[zp], P[zp], Pm1[i], cdt_dz);
            }
            for (int iz = Nz - 3; iz <= Nz - 1; ++iz) {  // z-max, innermost first
                const size_t i = IDX(ix, iy, iz, Nxy), zm = IDX(ix, iy, iz - 1, Nxy);
                P[i] = mur(Pm1[zm], P[zm], Pm1[i], cdt_dz_t);
            }
        }
    // x faces
    for (int iz = 0; iz < Nz; ++iz) {
        const double cdt_dxy = (p->z_min + iz * p->dz < p->water_tissue) ? p->c_inf * dt / p->dxy
                                                                          : p->c_inf_tissue * dt / p->dxy;
        (void)itiss;
        for (int iy = 0; iy < Nxy; ++iy) {
            for (int ix = Nxy - 3; ix <= Nxy - 1; ++ix) {
                const size_t i = IDX(ix, iy, iz, Nxy), xm = IDX(ix - 1, iy, iz, Nxy);
                P[i] = mur(Pm1[xm], P[xm], Pm1[i], cdt_dxy);
            }
            if (p->full_domain)
                for (int ix = 2; ix >= 0; --ix) {
                    const size_t i = IDX(ix, iy, iz, Nxy), xp = IDX(ix + 1, iy, iz, Nxy);
                    P[i] = mur(Pm1[xp], P[xp], Pm1[i], cdt_dxy);
                }
        }
    }
    // y faces
    for (int iz = 0; iz < Nz; ++iz) {
        const double cdt_dxy = (p->z_min + iz * p->dz < p->water_tissue) ? p->c_inf * dt / p->dxy
                                                                          : p->c_inf_tissue * dt / p->dxy;
        for (int ix = 0; ix < Nxy; ++ix) {
            for (int iy = Nxy - 3; iy <= Nxy - 1; ++iy) {
                const size_t i = IDX(ix, iy, iz, Nxy), ym = IDX(ix, iy - 1, iz, Nxy);
                P[i] = mur(Pm1[ym], P[ym], Pm1[i], cdt_dxy);
            }
            if (p->full_domain)
                for (int iy = 2; iy >= 0; --iy) {
                    const size_t i = IDX(ix, iy, iz, Nxy), yp = IDX(ix, iy + 1, iz, Nxy);
                    P[i] = mur(Pm1[yp], P[yp], Pm1[i], cdt_dxy);
                }
        }
    }
}

// W:242-265
void orc_acoustic_power_water(const bd_westervelt_params* p, double* Q, const double* Pnext,
                              const double* Pm1, int Nxy, int Nz, int last) {
    for (int iz = 0; iz < Nz; ++iz) {
        const double Z = p->z_min + iz * p->dz;
        if (!(Z < p->water_tissue)) continue;
        for (int iy = 0; iy < Nxy; ++iy)
            for (int ix = 0; ix < Nxy; ++ix) {
                const size_t idx = IDX(ix, iy, iz, Nxy);
                const double dP_dt_1 = (Pnext[idx] - Pm1[idx]) / (2.0 * p->dt);
                if (!last)
                    Q[idx] = Q[idx] + std::pow(dP_dt_1, 2) * p->dt;
                else
                    Q[idx] = 2.0 * p->abs_water * p->fr * Q[idx] /
                             (std::pow(2.0 * M_PI * p->fr, 2) * p->rho * p->c_inf);
            }
    }
}

// Running DFT bins k = n*periods of an N_time-sample record (replaces
// W:306-321 copy_data and the FFT of W:740-747).
void orc_harmonics_accumulate(double* H, const double* Pnext, int Nxy, int Nz, int iz_tissue,
                              int i_time, int N_time, int periods, int N_harm) {
    const size_t plane = (size_t)Nxy * Nxy, nf = plane * (size_t)(Nz - iz_tissue);
    for (int n = 1; n <= N_harm; ++n) {
        const long long m = ((long long)n * periods * i_time) % N_time;
        const double ang = 2.0 * M_PI * (double)m / (double)N_time;
        const double cs = std::cos(ang), sn = std::sin(ang);
        double* Hr = H + (size_t)(2 * (n - 1)) * nf;
        double* Hi = Hr + nf;
        for (size_t j = 0; j < nf; ++j) {
            const double v = Pnext[(size_t)iz_tissue * plane + j];
            Hr[j] += v * cs;
            Hi[j] -= v * sn;
        }
    }
}

// W:324-345 over the full tissue slab.
void orc_harmonics_q(const bd_westervelt_params* p, double* Q, const double* H, int Nxy, int Nz,
                     int iz_tissue, int N_time, int N_harm, double abs_coef) {
    const size_t plane = (size_t)Nxy * Nxy, nf = plane * (size_t)(Nz - iz_tissue);
    for (size_t j = 0; j < nf; ++j) {
        double q = 0.0;
        for (int n = 1; n <= N_harm; ++n) {
            const double re = H[(size_t)(2 * (n - 1)) * nf + j], im = H[(size_t)(2 * (n - 1) + 1) * nf + j];
            const double a = std::sqrt(re * re + im * im) / N_time;       // Abs_Complex, W:299
            const double In = 2.0 * std::pow(a, 2) / (p->rho_tissue * p->c_inf_tissue);
            q = q + (2.0 * (abs_coef * n * p->fr / std::pow(10.0, 6)) * In);
        }
        Q[(size_t)iz_tissue * plane + j] = q;
    }
}

size_t orc_wv_state_elems(int Nxy, int Nz) { return (size_t)12 * Nxy * Nxy * Nz; }

// Run n_steps steps starting at step index step0 on a 12-field state laid out
// as the library's plan: P ring slots 0..5 (p^m in slot m mod 6), R1 slots 6..8,
// R2 slots 9..11 (R^m in slot m mod 3).  Times follow the reference's repeated
// addition t += dt (W:564).  energy (2 per step), focus (1 per step) indexed
// by absolute step; both optional.
int orc_wv_run(const bd_westervelt_params* p, int Nxy, int Nz, int step0, int n_steps,
               double* state, double* pmax, double* energy, double* focus, int fx, int fy,
               int fz, const double* amp, const double* phase, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const size_t N = (size_t)Nxy * Nxy * Nz;
    auto Pslot = [&](int m) { return state + (size_t)(((m % 6) + 6) % 6) * N; };
    auto R1slot = [&](int m) { return state + (size_t)(6 + ((m % 3) + 3) % 3) * N; };
    auto R2slot = [&](int m) { return state + (size_t)(9 + ((m % 3) + 3) % 3) * N; };
    std::vector<double> e_plane(energy ? 2 * (size_t)Nz : 0, 0.0);
    double t = 0.0;
    for (int k = 0; k < step0 + 1; ++k) t = (k == 0) ? 0.0 : t + p->dt;  // t^{step0}
    for (int n = step0; n < step0 + n_steps; ++n) {
        double* Pnext = Pslot(n + 1);
        std::fill(e_plane.begin(), e_plane.end(), 0.0);
        wv_space6_impl(p, Pnext, Pslot(n), Pslot(n - 1), Pslot(n - 2), Pslot(n - 3), Pslot(n - 4),
                       R1slot(n), R1slot(n - 1), R1slot(n - 2), R2slot(n), R2slot(n - 1),
                       R2slot(n - 2), Nxy, Nz, pmax, energy ? e_plane.data() : nullptr);
        t = t + p->dt;  // t^{n+1}
        orc_wv_source(p, Pnext, t, Nxy, Nz, amp, phase);
        orc_wv_abc(p, Pnext, Pslot(n), t, Nxy, Nz);
        if (pmax) {  // boundary cells
            const int lo = p->full_domain ? 3 : 0;
            for (int iz = 0; iz < Nz; ++iz)
                for (int iy = 0; iy < Nxy; ++iy)
                    for (int ix = 0; ix < Nxy; ++ix) {
                        const bool interior = ix >= lo && ix <= Nxy - 4 && iy >= lo && iy <= Nxy - 4 &&
                                              iz >= 3 && iz <= Nz - 4;
                        if (interior) continue;
                        const size_t i = IDX(ix, iy, iz, Nxy);
                        if (Pnext[i] > pmax[i]) pmax[i] = Pnext[i];
                    }
        }
        if (energy) {
            double s0 = 0.0, s1 = 0.0;
            for (int iz = 0; iz < Nz; ++iz) { s0 += e_plane[2 * iz]; s1 += e_plane[2 * iz + 1]; }
            energy[2 * (size_t)n] = s0;
            energy[2 * (size_t)n + 1] = s1;
        }
        if (focus && fx >= 0) focus[n] = Pnext[IDX(fx, fy, fz, Nxy)];
    }
    return 0;
}

}  // extern "C"
