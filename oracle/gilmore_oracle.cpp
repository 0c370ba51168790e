// FP64 serial restatement of the reference's Gilmore bubble model.
//
// TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU
// baseline; the shipped library never links or calls it.
//
// Reference: Gilmore_liver_3fr_combo.cu (cited as G:line).
//   drive P, P_dt                  G:40-78
//   P_in, dP_in_dt                 G:82-92
//   LocalSoundSpeed                G:96-103
//   Stress_dt, Q_dt                G:106-117
//   Enthalpy, Enthalpy_dt          G:121-157
//   F_R (Gilmore equation)         G:161-179
//   norm_Euclid                    G:183-192
//   DPRK_step (DP5(4) + control)   G:212-318
//   threshold search kernel        G:323-429
//   frequency table                G:465-469, G:498-520
// Expressions keep the reference's operation order; compiled without FMA
// contraction or fast-math.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/bubbledynamics.h"

namespace {

// General drive sum_k A_k sin(2 pi f_k t + phi_k), or the reference's
// single/dual/triple form when `ref` is set.
struct Drive {
    bool ref;
    bool triple;                   // reference triple_fr switch
    double A;                      // reference amplitude
    double f[3], a[3], ph[3];
};

double P(const Drive& d, double t) {
    if (d.ref) {  // G:40-57
        const double fr1 = d.f[0], fr2 = d.f[1], fr3 = d.f[2], Amplitude = d.A;
        if (!d.triple) return Amplitude * std::sin(2.0 * M_PI * fr1 * t);
        if (fr1 == fr2 && fr1 == fr3) return Amplitude * std::sin(2.0 * M_PI * fr1 * t);
        if (fr1 == fr2 || fr2 == fr3)
            return (Amplitude / std::sqrt(2.0)) * (std::sin(2.0 * M_PI * fr1 * t) + std::sin(2.0 * M_PI * fr3 * t));
        return (Amplitude / std::sqrt(3.0)) *
               (std::sin(2.0 * M_PI * fr1 * t) + std::sin(2.0 * M_PI * fr2 * t) + std::sin(2.0 * M_PI * fr3 * t));
    }
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += d.a[k] * std::sin(2.0 * M_PI * d.f[k] * t + d.ph[k]);
    return s;
}

double P_dt(const Drive& d, double t) {
    if (d.ref) {  // G:61-78
        const double fr1 = d.f[0], fr2 = d.f[1], fr3 = d.f[2], Amplitude = d.A;
        if (!d.triple) return Amplitude * 2.0 * M_PI * fr1 * std::cos(2.0 * M_PI * fr1 * t);
        if (fr1 == fr2 && fr1 == fr3) return Amplitude * 2.0 * M_PI * fr1 * std::cos(2.0 * M_PI * fr1 * t);
        if (fr1 == fr2 || fr2 == fr3)
            return Amplitude * std::sqrt(2.0) * M_PI *
                   (fr1 * std::cos(2.0 * M_PI * fr1 * t) + fr3 * std::cos(2.0 * M_PI * fr3 * t));
        return Amplitude * (2.0 / std::sqrt(3.0)) * M_PI *
               (fr1 * std::cos(2.0 * M_PI * fr1 * t) + fr2 * std::cos(2.0 * M_PI * fr2 * t) +
                fr3 * std::cos(2.0 * M_PI * fr3 * t));
    }
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += d.a[k] * 2.0 * M_PI * d.f[k] * std::cos(2.0 * M_PI * d.f[k] * t + d.ph[k]);
    return s;
}

struct Model {
    const bd_gilmore_params* g;
    Drive d;
    double R0;
};

double P_in(const Model& m, double R) {  // G:82-85
    return (m.g->P0 + (m.g->S2 / m.R0)) * std::pow(m.R0 / R, m.g->gamma3);
}

double dP_in_dt(const Model& m, double R, double R1) {  // G:89-92
    return (m.g->P0 + (m.g->S2 / m.R0)) * std::pow(m.R0, m.g->gamma3) * (-m.g->gamma3) *
           std::pow(R, -m.g->gamma3 - 1.0) * R1;
}

double LocalSoundSpeed(const Model& m, double t, double R, double Tau) {  // G:96-103
    const double p = P(m.d, t);
    const double p_in = P_in(m, R);
    double C = ((p_in - (m.g->S2 / R) + Tau) + m.g->B) / (m.g->P0 + p + m.g->B);
    C = m.g->C_inf * std::pow(C, (m.g->n - 1.0) / (2.0 * m.g->n));
    return C;
}

double Stress_dt(const Model& m, double R, double R1, double Tau) {  // G:106-110
    const double Tau_dt = -m.g->G4_3 * (1.0 - std::pow(m.R0 / R, 3)) - (m.g->mu4 * R1 / R) - Tau;
    return Tau_dt / m.g->lambda;
}

double Q_dt(const Model& m, double R, double R1, double Tau, double q) {  // G:113-117
    const double q_dt = (-m.g->G4_3 * (1.0 - std::pow(m.R0 / R, 3)) - (m.g->mu4 * R1 / R)) / (3.0 * m.g->lambda);
    return q_dt - (q / m.g->lambda) - (R1 * Tau / R);
}

double Enthalpy(const Model& m, double t, double R, double Tau) {  // G:121-128
    const double p = P(m.d, t);
    const double p_in = P_in(m, R);
    const double part1 = (m.g->P0 + p + m.g->B) * m.g->n / (m.g->rho * (m.g->n - 1.0));
    const double part2 = (p_in - (m.g->S2 / R) + Tau + m.g->B) / (m.g->P0 + p + m.g->B);
    return part1 * (std::pow(part2, (m.g->n - 1.0) / m.g->n) - 1.0);
}

double Enthalpy_dt(const Model& m, double t, double R, double R1, double Tau, double Tau1) {  // G:132-157
    const double n = m.g->n;
    const double p = P(m.d, t);
    const double p_in = P_in(m, R);
    const double part1 = (m.g->P0 + p + m.g->B) * n / (m.g->rho * (n - 1.0));
    const double p_dt = P_dt(m.d, t);
    const double part1_dt = p_dt * n / (m.g->rho * (n - 1.0));
    const double u = p_in - (m.g->S2 / R) + Tau + m.g->B;
    const double v = m.g->P0 + p + m.g->B;
    const double I = u / v;
    const double part2 = std::pow(I, (n - 1.0) / n) - 1.0;
    const double d_pin = dP_in_dt(m, R, R1);
    const double u_dt = d_pin + (m.g->S2 * R1 / (R * R)) + Tau1;
    const double v_dt = p_dt;
    const double I_dt = (u_dt * v - u * v_dt) / std::pow(v, 2);
    const double part2_dt = ((n - 1.0) / n) * std::pow(I, -1.0 / n) * I_dt;
    return part1_dt * part2 + part1 * part2_dt;
}

double F_R(const Model& m, double t, double R, double R1, double Tau, double q) {  // G:161-179
    const double Tau1 = Stress_dt(m, R, R1, Tau);
    const double q1 = Q_dt(m, R, R1, Tau, q);
    const double H = Enthalpy(m, t, R, Tau);
    const double H1 = Enthalpy_dt(m, t, R, R1, Tau, Tau1);
    const double C = LocalSoundSpeed(m, t, R, Tau);
    const double a = 1.0 - (R1 / C);
    const double b = 1.0 + (R1 / C);
    const double part1 = 1.5 * std::pow(R1, 2) * (1.0 - (R1 / (3.0 * C)));
    const double part2 = b * (H + ((-Tau + 3.0 * q) / m.g->rho));
    const double part3 = (H1 * a + ((-Tau1 + 3.0 * q1) / m.g->rho)) * R / C;
    return (-part1 + part2 + part3) / (a * R);
}

double norm_Euclid(const double* error, const double* sc, int length) {  // G:183-192
    double S = 0.0;
    for (int i = 0; i < length; i++) S = S + std::pow(error[i] / sc[i], 2);
    S = S / length;
    return std::sqrt(S);
}

// Butcher tableau of DP5(4), G:218-227 (b: 5th order, propagated; bb: 4th order).
// The reference writes a32 = -9/40 (G:220); Dormand & Prince (1980) publish
// +9/40.  Both are kept: tableau 0 reproduces the reference, 1 the published method.
const double A_[7][7] = {{0, 0, 0, 0, 0, 0, 0},
                         {0.2, 0, 0, 0, 0, 0, 0},
                         {3.0 / 40.0, -9.0 / 40.0, 0, 0, 0, 0, 0},
                         {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
                         {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
                         {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
                         {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
const double B_[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
const double BB_[7] = {5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
const double C_[7] = {0.0, 0.2, 0.3, 0.8, 8.0 / 9.0, 1.0, 1.0};

inline double a_coef(int tab, int i, int j) {
    if (tab == BD_GM_TAB_DP1980 && i == 2 && j == 1) return 9.0 / 40.0;
    return A_[i][j];
}

// G:212-318.  Returns the number of rejected attempts.
int DPRK_step(const Model& m, double& t, double& R, double& R1, double& Tau, double& q, double& h,
              bool& threshold, double endT, double abs_eps, double rel_eps) {
    const int tab = m.g->tableau;
    double Y[4], error[4], sc[4];
    double K[7], L[7], Tau_K[7], q_K[7];
    double k_R, k_R1, k_Tau, k_q, err, max_result = 0.0;
    int rejected = 0;
    if (t + h > endT) h = endT - t;
    err = 2.0;
    double fac_max = 5.0;
    threshold = false;
    while (err > 1.0) {
        Y[0] = R; Y[1] = R1; Y[2] = Tau; Y[3] = q;
        error[0] = error[1] = error[2] = error[3] = 0.0;
        for (int i = 0; i < 7; i++) K[i] = L[i] = Tau_K[i] = q_K[i] = 0.0;
        for (int i = 0; i < 7; i++) {
            k_R = 0.0; k_R1 = 0.0; k_Tau = 0.0; k_q = 0.0;
            for (int j = 0; j < i; j++) {
                const double aij = a_coef(tab, i, j);
                k_R = k_R + aij * K[j] * h;
                k_R1 = k_R1 + aij * L[j] * h;
                k_Tau = k_Tau + aij * Tau_K[j] * h;
                k_q = k_q + aij * q_K[j] * h;
            }
            K[i] = (R1 + k_R1);
            L[i] = F_R(m, t + h * C_[i], R + k_R, R1 + k_R1, Tau + k_Tau, q + k_q);
            Tau_K[i] = Stress_dt(m, R + k_R, R1 + k_R1, Tau + k_Tau);
            q_K[i] = Q_dt(m, R + k_R, R1 + k_R1, Tau + k_Tau, q + k_q);
            Y[0] = Y[0] + B_[i] * K[i] * h;
            Y[1] = Y[1] + B_[i] * L[i] * h;
            Y[2] = Y[2] + B_[i] * Tau_K[i] * h;
            Y[3] = Y[3] + B_[i] * q_K[i] * h;
            error[0] = error[0] + K[i] * h * (B_[i] - BB_[i]);
            error[1] = error[1] + L[i] * h * (B_[i] - BB_[i]);
            error[2] = error[2] + Tau_K[i] * h * (B_[i] - BB_[i]);
            error[3] = error[3] + q_K[i] * h * (B_[i] - BB_[i]);
        }
        sc[0] = abs_eps + rel_eps * std::fmax(std::fabs(R), std::fabs(Y[0]));
        sc[1] = abs_eps + rel_eps * std::fmax(std::fabs(R1), std::fabs(Y[1]));
        sc[2] = abs_eps + rel_eps * std::fmax(std::fabs(Tau), std::fabs(Y[2]));
        sc[3] = abs_eps + rel_eps * std::fmax(std::fabs(q), std::fabs(Y[3]));
        err = norm_Euclid(error, sc, 4);
        max_result = std::fmax(0.1, 0.9 * std::pow(err, -0.2));
        if (err > 1.0) {
            fac_max = 0.9;
            h = h * std::fmin(fac_max, max_result);
            ++rejected;
            if (!(h > 0.0) || !std::isfinite(err)) break;  // guard: a NaN state never satisfies err <= 1
        }
    }
    t = t + h;
    R = Y[0]; R1 = Y[1]; Tau = Y[2]; q = Y[3];
    h = h * std::fmin(fac_max, max_result);
    if (Y[1] <= -m.g->C_inf) threshold = true;
    return rejected;
}

// Classical RK4 step of the same system.
void rk4_step(const Model& m, double t, double h, double y[4]) {
    auto f = [&](double tt, const double s[4], double o[4]) {
        o[0] = s[1];
        o[1] = F_R(m, tt, s[0], s[1], s[2], s[3]);
        o[2] = Stress_dt(m, s[0], s[1], s[2]);
        o[3] = Q_dt(m, s[0], s[1], s[2], s[3]);
    };
    double k1[4], k2[4], k3[4], k4[4], s[4];
    f(t, y, k1);
    for (int i = 0; i < 4; ++i) s[i] = y[i] + 0.5 * h * k1[i];
    f(t + 0.5 * h, s, k2);
    for (int i = 0; i < 4; ++i) s[i] = y[i] + 0.5 * h * k2[i];
    f(t + 0.5 * h, s, k3);
    for (int i = 0; i < 4; ++i) s[i] = y[i] + h * k3[i];
    f(t + h, s, k4);
    for (int i = 0; i < 4; ++i) y[i] = y[i] + h / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

}  // namespace

extern "C" {

void orc_gilmore_reference_params(bd_gilmore_params* p) {  // G:440-460
    std::memset(p, 0, sizeof(*p));
    p->rho = 1100.0;
    p->P0 = 1.013 * std::pow(10.0, 5);
    p->S2 = 2.0 * 0.056;
    p->mu4 = 4.0 * 0.009;
    p->G4_3 = 0.04 * std::pow(10.0, 6) * 4.0 / 3.0;
    p->C_inf = 1549.0;
    p->lambda = 3.0 * std::pow(10.0, -9);
    p->n = 7.0;
    p->B = (p->C_inf * p->C_inf * p->rho / p->n) - p->P0;
    p->gamma1 = 1.4;
    p->gamma3 = 3.0 * p->gamma1;
    p->triple_fr = 1;
    p->tableau = BD_GM_TAB_REFERENCE;
    p->minA = std::pow(10.0, 4);
    p->stepA = 1000.0;
    p->stepR0 = std::pow(10.0, -7);
    p->minR0 = std::pow(10.0, -7);
    p->abs_eps = std::pow(10.0, -9);
    p->rel_eps = std::pow(10.0, -9);
    p->h0 = std::pow(10.0, -13);
    p->max_solve_steps = 1e7;
    p->arithmetic = BD_GM_ARITH_REFERENCE;
}

void orc_dp_tableau(int tab, double* a, double* b, double* bb, double* c) {
    for (int i = 0; i < 7; ++i) {
        for (int j = 0; j < 7; ++j) a[7 * i + j] = a_coef(tab, i, j);
        b[i] = B_[i]; bb[i] = BB_[i]; c[i] = C_[i];
    }
}

// Frequency triplets of G:465-469 and G:498-520.  Writes 3*count doubles
// (pass out = NULL to count) and returns the number of combos.
int orc_gilmore_freq_table(double min_fr, double max_fr, int N_fr, double* out) {
    const double step_log = (std::log10(max_fr) - std::log10(min_fr)) / (N_fr - 1.0);
    int i_fr_combo = 0;
    for (int i = 0; i < N_fr; i++) {
        bool first_cycle = true;
        for (int ii = i; ii < N_fr; ii++) {
            int iii = ii;
            if (!first_cycle) iii = ii + 1;
            for (; iii < N_fr; iii++) {
                if (out) {
                    out[3 * i_fr_combo + 0] = std::pow(10.0, std::log10(min_fr) + i * step_log);
                    out[3 * i_fr_combo + 1] = std::pow(10.0, std::log10(min_fr) + ii * step_log);
                    out[3 * i_fr_combo + 2] = std::pow(10.0, std::log10(min_fr) + iii * step_log);
                }
                i_fr_combo++;
            }
            first_cycle = false;
        }
    }
    return i_fr_combo;
}

// Right-hand side of the reference model at one state, general drive
// (amp/phase/freq 3 each) or, with amp == NULL, the reference drive of
// amplitude A_ref and frequencies freq.  out = (dR/dt, d2R/dt2, dTau/dt, dq/dt)
// plus (H, dH/dt, C) in out[4..6].
void orc_gilmore_rhs(const bd_gilmore_params* g, double R0, const double* amp, const double* phase,
                     const double* freq, double A_ref, double t, const double* y, double* out) {
    Model m{g, {}, R0};
    m.d.ref = amp == nullptr;
    m.d.triple = g->triple_fr != 0;
    m.d.A = A_ref;
    for (int k = 0; k < 3; ++k) {
        m.d.f[k] = freq[k];
        m.d.a[k] = amp ? amp[k] : 0.0;
        m.d.ph[k] = phase ? phase[k] : 0.0;
    }
    out[0] = y[1];
    out[1] = F_R(m, t, y[0], y[1], y[2], y[3]);
    out[2] = Stress_dt(m, y[0], y[1], y[2]);
    out[3] = Q_dt(m, y[0], y[1], y[2], y[3]);
    out[4] = Enthalpy(m, t, y[0], y[2]);
    out[5] = Enthalpy_dt(m, t, y[0], y[1], y[2], out[2]);
    out[6] = LocalSoundSpeed(m, t, y[0], y[2]);
}

// One DP5(4) step of the reference (G:212-318) from state y at t with step h
// (reference drive A_ref, freq).  Updates t, y, h; returns rejections and
// sets *threshold.
int orc_gilmore_dprk_step(const bd_gilmore_params* g, double R0, const double* freq, double A_ref,
                          double endT, double* t, double* y, double* h, int* threshold) {
    Model m{g, {}, R0};
    m.d.ref = true;
    m.d.triple = g->triple_fr != 0;
    m.d.A = A_ref;
    for (int k = 0; k < 3; ++k) m.d.f[k] = freq[k];
    bool thr = false;
    const int rej = DPRK_step(m, *t, y[0], y[1], y[2], y[3], *h, thr, endT, g->abs_eps, g->rel_eps);
    *threshold = thr ? 1 : 0;
    return rej;
}

// One ODE solve of the threshold search at amplitude A from the rest state
// with first step h0 (the inner loops of G:378-379, G:398-399, G:418-419).
// Returns 1 on collapse, 0 otherwise; *steps = accepted steps.
int orc_gilmore_solve(const bd_gilmore_params* g, double R0, const double* freq, double A, double endT,
                      double h0, long long* steps) {
    Model m{g, {}, R0};
    m.d.ref = true;
    m.d.triple = g->triple_fr != 0;
    m.d.A = A;
    for (int k = 0; k < 3; ++k) m.d.f[k] = freq[k];
    const long long cap = g->max_solve_steps > 0 ? (long long)g->max_solve_steps : 10000000LL;
    double t = 0.0, R = R0, R1 = 0.0, Tau = 0.0, q = 0.0, h = h0;
    bool thr = false;
    long long n = 0;
    while (t < endT && !thr && n++ < cap)
        DPRK_step(m, t, R, R1, Tau, q, h, thr, endT, g->abs_eps, g->rel_eps);
    if (steps) *steps = n;
    return thr ? 1 : 0;
}

// G:323-429 for thread indices [idx0, idx1); results_A indexed by idx.
void orc_gilmore_threshold(const bd_gilmore_params* g, double* results_A, const double* Fr_combo,
                           int N_fr_combo, int N_R0, int idx0, int idx1, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int idx = idx0; idx < idx1; ++idx) {
        const int i_fr_combo = idx / N_R0;
        const int i_R0 = idx % N_R0;
        if (!(i_fr_combo >= 0 && i_fr_combo < N_fr_combo && i_R0 >= 0 && i_R0 < N_R0)) continue;
        const double R0 = g->minR0 + i_R0 * g->stepR0;
        const double fr1 = Fr_combo[3 * i_fr_combo], fr2 = Fr_combo[3 * i_fr_combo + 1],
                     fr3 = Fr_combo[3 * i_fr_combo + 2];
        double endT = 0.0;  // G:342-353
        if (g->triple_fr) {
            if (fr1 <= fr2 && fr1 <= fr3) endT = 5.0 / fr1;
            if (fr2 <= fr1 && fr2 <= fr3) endT = 5.0 / fr2;
            if (fr3 <= fr1 && fr3 <= fr2) endT = 5.0 / fr3;
        } else {
            endT = 5.0 / fr1;
        }
        double h = g->h0;
        results_A[idx] = -1.0;
        double t = 0.0, R = R0, R1 = 0.0, Tau = 0.0, q = 0.0;
        bool threshold_5 = false, threshold_4 = true, threshold_3 = true;
        double Amplitude_5 = 0.0, Amplitude_4 = 0.0, Amplitude_3 = 0.0;
        Model m{g, {}, R0};
        m.d.ref = true;
        m.d.triple = g->triple_fr != 0;
        m.d.f[0] = fr1; m.d.f[1] = fr2; m.d.f[2] = fr3;
        const long long cap = g->max_solve_steps > 0 ? (long long)g->max_solve_steps : 10000000LL;
        long long steps = 0;
        // A solve that ends neither in collapse nor at endT within the cap
        // leaves the search undecided: the cell's result is BD_GM_SWEEP_CAPPED
        // (the reference has no cap and would keep integrating).
        bool capped = false;
        // (1) 0.1 MPa steps; h carries over between amplitudes (G:374)
        while (!threshold_5 && Amplitude_5 < std::pow(10.0, 7)) {
            t = 0.0; R = R0; R1 = 0.0; Tau = 0.0; q = 0.0;
            Amplitude_5 = Amplitude_5 + std::pow(10.0, 5);
            threshold_5 = false;
            m.d.A = Amplitude_5;
            steps = 0;
            while (t < endT && !threshold_5 && steps++ < cap)
                DPRK_step(m, t, R, R1, Tau, q, h, threshold_5, endT, g->abs_eps, g->rel_eps);
            if (!threshold_5 && t < endT) { capped = true; break; }
        }
        if (capped) { results_A[idx] = BD_GM_SWEEP_CAPPED; continue; }
        if (!threshold_5 && Amplitude_5 == std::pow(10.0, 7)) results_A[idx] = Amplitude_5;
        // (2) 0.01 MPa steps down
        Amplitude_4 = Amplitude_5;
        while (threshold_5 && threshold_4) {
            h = g->h0;
            t = 0.0; R = R0; R1 = 0.0; Tau = 0.0; q = 0.0;
            Amplitude_4 = Amplitude_4 - std::pow(10.0, 4);
            threshold_4 = false;
            m.d.A = Amplitude_4;
            steps = 0;
            while (t < endT && !threshold_4 && steps++ < cap)
                DPRK_step(m, t, R, R1, Tau, q, h, threshold_4, endT, g->abs_eps, g->rel_eps);
            if (!threshold_4 && t < endT) { capped = true; break; }
            if (!threshold_4) Amplitude_4 = Amplitude_4 + std::pow(10.0, 4);
        }
        if (capped) { results_A[idx] = BD_GM_SWEEP_CAPPED; continue; }
        // (3) 0.001 MPa steps down
        Amplitude_3 = Amplitude_4;
        while (threshold_5 && !threshold_4 && threshold_3) {
            h = g->h0;
            t = 0.0; R = R0; R1 = 0.0; Tau = 0.0; q = 0.0;
            Amplitude_3 = Amplitude_3 - std::pow(10.0, 3);
            threshold_3 = false;
            m.d.A = Amplitude_3;
            steps = 0;
            while (t < endT && !threshold_3 && steps++ < cap)
                DPRK_step(m, t, R, R1, Tau, q, h, threshold_3, endT, g->abs_eps, g->rel_eps);
            if (!threshold_3 && t < endT) { capped = true; break; }
            if (!threshold_3) results_A[idx] = Amplitude_3 + std::pow(10.0, 3);
        }
        if (capped) results_A[idx] = BD_GM_SWEEP_CAPPED;
    }
}

// Batch integration with the general drive (the bd_gilmore_batch semantics):
// RK4 (method 0) or DP5(4) with rtol/atol (method 1).
void orc_gilmore_integrate(const bd_gilmore_params* g, int method, int n, const double* R0,
                           const double* amp, const double* phase, const double* freq, double t_end,
                           double dt, double rtol, double atol, int out_every, int n_out,
                           long long max_steps, double* R_out, double* Rmax, double* t_collapse,
                           long long* n_acc, long long* n_rej, int i0, int i1, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    (void)n;
    if (max_steps <= 0) max_steps = 10000000LL;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = i0; i < i1; ++i) {
        Model m{g, {}, R0[i]};
        m.d.ref = false;
        for (int k = 0; k < 3; ++k) {
            m.d.f[k] = freq[k];
            m.d.a[k] = amp[3 * i + k];
            m.d.ph[k] = phase[3 * i + k];
        }
        double y[4] = {R0[i], 0.0, 0.0, 0.0};
        double rmax = y[0], tc = -1.0, t = 0.0;
        long long acc = 0, rej = 0;
        int j_out = 0;
        if (R_out && n_out > 0) R_out[(size_t)i * n_out + j_out++] = y[0];
        if (method == BD_GM_RK4) {
            const long long nsteps = (long long)std::llround(t_end / dt);
            for (long long s = 0; s < nsteps; ++s) {
                if (max_steps > 0 && acc >= max_steps) break;
                rk4_step(m, s * dt, dt, y);
                ++acc;
                t = (s + 1) * dt;
                if (y[0] > rmax) rmax = y[0];
                if (R_out && out_every > 0 && (s + 1) % out_every == 0 && j_out < n_out)
                    R_out[(size_t)i * n_out + j_out++] = y[0];
                if (y[1] <= -g->C_inf) { tc = t; break; }
            }
        } else {
            double h = dt;
            bool thr = false;
            while (t < t_end && !thr) {
                if (max_steps > 0 && acc + rej >= max_steps) break;
                rej += DPRK_step(m, t, y[0], y[1], y[2], y[3], h, thr, t_end, atol, rtol);
                ++acc;
                if (y[0] > rmax) rmax = y[0];
                if (!(h > 0.0) || !std::isfinite(y[0])) break;
            }
            if (thr) tc = t;
        }
        if (R_out) for (; j_out < n_out; ++j_out) R_out[(size_t)i * n_out + j_out] = NAN;
        Rmax[i] = rmax;
        t_collapse[i] = tc;
        if (n_acc) n_acc[i] = acc;
        if (n_rej) n_rej[i] = rej;
    }
}

}  // extern "C"
