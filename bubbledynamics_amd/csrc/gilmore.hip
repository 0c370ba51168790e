// Gilmore bubble dynamics in viscoelastic tissue on gfx950.
//
// Reference: Gilmore_liver_3fr_combo.cu (G:line).
//
// Two arithmetics of the right-hand side (G:40-179):
//  * reference: the reference's expressions in its evaluation order (the
//    cancelling pow(I, (n-1)/n) - 1 and (p_in - 2s/R + tau + B)/(P0 + p + B)
//    included), with every step's first stage recomputed (G:259-273).  Only
//    the libm rounding of pow/sin/cos is left open against the reference;
//    threshold-sweep results are compared exactly with the FP64 oracle.
//  * fast: common subexpressions shared,
//      p, dp/dt        one sincos per tone
//      p_in, dp_in/dt  p_in = (P0 + 2s/R0) exp(3g log(R0/R)),  dp_in/dt = -3g p_in R'/R
//      I^((n-1)/n) - 1 = expm1(((n-1)/n) log1p((u - v)/v)), u - v = p_in - 2s/R + tau - P0 - p
//                      (B cancels exactly, so nothing cancels when I ~ 1)
//      C = c_inf sqrt(I^((n-1)/n)),  I^(-1/n) = I^((n-1)/n) / I
//    and the last stage of a DP5(4) step reused as the next first stage
//    (FSAL): 6 right-hand sides per attempt instead of 7.
// Integrators: DP5(4) with the step controller of G:212-318 and classical
// RK4.  Kernels are persistent: every lane advances its own bubble (or
// threshold-search cell) by one Runge-Kutta attempt per loop trip and takes
// the next one from a device counter as soon as it finishes, so uneven step
// counts do not idle lanes.  State, error norm and controller are FP64; with
// BD_GM_SPLIT the fast stage right-hand sides run in FP32 while R/R0 >= split_ratio.
#include "common.h"
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

// Scratch of bd_gilmore_threshold_sweep_noalloc: the sweep's queue counters
// (5 used), one copy per GPU for the life of the process.  In host code this
// name is a host-side stand-in: kernels get the device address that
// hipGetSymbolAddress returns for it (noalloc_scratch_ptr below), never the name.
static __device__ __align__(256) int scratch[64];

namespace bd {

template <typename U>
struct GmModel {
    U rho, inv_rho, P0, S2, B, C_inf, inv_lambda, inv_3lambda, nm1_n, n_rho_nm1, gamma3, mu4, G4_3;
};

template <typename U>
static GmModel<U> make_model(const bd_gilmore_params* g) {
    GmModel<U> m;
    m.rho = (U)g->rho;
    m.inv_rho = (U)(1.0 / g->rho);
    m.P0 = (U)g->P0;
    m.S2 = (U)g->S2;
    m.B = (U)g->B;
    m.C_inf = (U)g->C_inf;
    m.inv_lambda = (U)(1.0 / g->lambda);
    m.inv_3lambda = (U)(1.0 / (3.0 * g->lambda));
    m.nm1_n = (U)((g->n - 1.0) / g->n);
    m.n_rho_nm1 = (U)(g->n / (g->rho * (g->n - 1.0)));
    m.gamma3 = (U)g->gamma3;
    m.mu4 = (U)g->mu4;
    m.G4_3 = (U)g->G4_3;
    return m;
}

// The reference's constants as its expressions use them (G:13-23).
struct GmRefModel {
    double rho, P0, S2, C_inf, lambda, n, B, gamma3, mu4, G4_3;
};

static GmRefModel make_ref_model(const bd_gilmore_params* g) {
    GmRefModel m;
    m.rho = g->rho; m.P0 = g->P0; m.S2 = g->S2; m.C_inf = g->C_inf; m.lambda = g->lambda;
    m.n = g->n; m.B = g->B; m.gamma3 = g->gamma3; m.mu4 = g->mu4; m.G4_3 = g->G4_3;
    return m;
}

// DP5(4) tableau (G:218-227): a, b (5th order, propagated), e = b - bb, c.
struct GmTab {
    double a[7][7], b[7], e[7], c[7];
};

static GmTab make_tab(int which) {
    static const double A[7][7] = {
        {0, 0, 0, 0, 0, 0, 0},
        {0.2, 0, 0, 0, 0, 0, 0},
        {3.0 / 40.0, -9.0 / 40.0, 0, 0, 0, 0, 0},
        {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
        {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
        {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
        {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
    static const double Bv[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
    static const double BB[7] = {5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0,
                                 187.0 / 2100.0, 1.0 / 40.0};
    static const double Cv[7] = {0.0, 0.2, 0.3, 0.8, 8.0 / 9.0, 1.0, 1.0};
    GmTab t;
    for (int i = 0; i < 7; ++i) {
        for (int j = 0; j < 7; ++j) t.a[i][j] = A[i][j];
        t.b[i] = Bv[i];
        t.e[i] = Bv[i] - BB[i];
        t.c[i] = Cv[i];
    }
    // the reference writes a32 = -9/40 (G:220); Dormand & Prince (1980) publish +9/40
    if (which == BD_GM_TAB_DP1980) t.a[2][1] = 9.0 / 40.0;
    return t;
}

struct GmConsts {
    GmModel<double> m64;
    GmModel<float> m32;
    GmRefModel ref;
    GmTab tab;
    double C_inf, abs_eps, rel_eps, h0;
    long long cap;
};

// p(t) = sum_k A[k] sin(w[k] t + ph[k])
struct GmDrive {
    double w[3], A[3], ph[3];
};

template <typename U>
__device__ __forceinline__ void drive(const GmDrive& d, double t, U& p, U& pd) {
    p = U(0);
    pd = U(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (d.A[k] == 0.0) continue;
        double ph = d.w[k] * t + d.ph[k];
        U s, c;
        if constexpr (std::is_same<U, float>::value) {
            // reduce in FP64, then an FP32 sincos on [-pi, pi]
            ph = ph - 6.283185307179586 * rint(ph * 0.15915494309189535);
            s = sinf((float)ph);
            c = cosf((float)ph);
        } else {
            sincos(ph, &s, &c);
        }
        p += (U)d.A[k] * s;
        pd += (U)(d.A[k] * d.w[k]) * c;
    }
}

// State derivative (R', R'', tau', q') of G:161-179 in arithmetic U, shared
// subexpressions (the fast arithmetic).
template <typename U>
__device__ __forceinline__ void gm_rhs(const GmModel<U>& m, const GmDrive& d, double t, U R0, U Pin0,
                                       const U y[4], U f[4]) {
    const U R = y[0], R1 = y[1], Tau = y[2], q = y[3];
    U p, pd;
    drive<U>(d, t, p, pd);
    const U inv_R = U(1) / R;
    const U x = R0 * inv_R;
    const U pin = Pin0 * exp(m.gamma3 * log(x));                          // G:82-85
    const U dpin = -m.gamma3 * pin * R1 * inv_R;                           // G:89-92
    const U x3 = x * x * x;
    const U el = -m.G4_3 * (U(1) - x3) - m.mu4 * R1 * inv_R;
    const U tau1 = (el - Tau) * m.inv_lambda;                              // G:106-110
    const U q1 = el * m.inv_3lambda - q * m.inv_lambda - R1 * Tau * inv_R;  // G:113-117
    const U v = m.P0 + p + m.B;                                            // G:146
    const U duv = pin - m.S2 * inv_R + Tau - m.P0 - p;                     // u - v
    const U w = duv / v;                                                   // I - 1
    const U L = log1p(w);
    const U part2 = expm1(m.nm1_n * L);                                    // I^((n-1)/n) - 1
    const U Ipow = U(1) + part2;
    const U C = m.C_inf * sqrt(Ipow);                                      // G:96-103
    const U part1 = v * m.n_rho_nm1;
    const U H = part1 * part2;                                             // G:121-128
    const U part1_dt = pd * m.n_rho_nm1;
    const U u_dt = dpin + m.S2 * R1 * inv_R * inv_R + tau1;
    const U I = U(1) + w;
    const U I_dt = (u_dt - I * pd) / v;
    const U part2_dt = m.nm1_n * (Ipow / I) * I_dt;
    const U H1 = part1_dt * part2 + part1 * part2_dt;                      // G:132-157
    const U inv_C = U(1) / C;
    const U rc = R1 * inv_C;
    const U a = U(1) - rc, b = U(1) + rc;
    const U t1 = U(1.5) * R1 * R1 * (U(1) - rc * U(1.0 / 3.0));
    const U t2 = b * (H + (U(3) * q - Tau) * m.inv_rho);
    const U t3 = (H1 * a + (U(3) * q1 - tau1) * m.inv_rho) * R * inv_C;
    f[0] = R1;
    f[1] = (-t1 + t2 + t3) / (a * R);                                      // G:161-179
    f[2] = tau1;
    f[3] = q1;
}

// RHS on the FP64 state, evaluated in U (fast arithmetic).
template <typename U>
__device__ __forceinline__ void rhs_at(const GmConsts& K, const GmDrive& d, double t, double R0, double Pin0,
                                       const double y[4], double f[4]) {
    U yu[4], fu[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) yu[c] = (U)y[c];
    if constexpr (std::is_same<U, float>::value)
        gm_rhs<float>(K.m32, d, t, (float)R0, (float)Pin0, yu, fu);
    else
        gm_rhs<double>(K.m64, d, t, R0, Pin0, yu, fu);
#pragma unroll
    for (int c = 0; c < 4; ++c) f[c] = (double)fu[c];
}

// ---------------------------------------------------------------------------
// reference arithmetic (G:40-179 in the reference's evaluation order)
// ---------------------------------------------------------------------------
// Drive: mode 0 one tone f[0] (G:42-47), 1 dual f[0] and f[2] with A/sqrt 2
// (G:51-52), 2 triple with A/sqrt 3 (G:55), 3 the batch's general drive
// sum_k a[k] sin(2 pi f[k] t + ph[k]).
struct GmRefDrive {
    int mode;
    double A;
    double f[3], a[3], ph[3];
};

__device__ __forceinline__ double ref_P(const GmRefDrive& d, double t) {            // G:40-57
    if (d.mode == 0) return d.A * sin(2.0 * M_PI * d.f[0] * t);
    if (d.mode == 1) return (d.A / sqrt(2.0)) * (sin(2.0 * M_PI * d.f[0] * t) + sin(2.0 * M_PI * d.f[2] * t));
    if (d.mode == 2)
        return (d.A / sqrt(3.0)) *
               (sin(2.0 * M_PI * d.f[0] * t) + sin(2.0 * M_PI * d.f[1] * t) + sin(2.0 * M_PI * d.f[2] * t));
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + d.a[k] * sin(2.0 * M_PI * d.f[k] * t + d.ph[k]);
    return s;
}

__device__ __forceinline__ double ref_P_dt(const GmRefDrive& d, double t) {         // G:61-78
    if (d.mode == 0) return d.A * 2.0 * M_PI * d.f[0] * cos(2.0 * M_PI * d.f[0] * t);
    if (d.mode == 1)
        return d.A * sqrt(2.0) * M_PI *
               (d.f[0] * cos(2.0 * M_PI * d.f[0] * t) + d.f[2] * cos(2.0 * M_PI * d.f[2] * t));
    if (d.mode == 2)
        return d.A * (2.0 / sqrt(3.0)) * M_PI *
               (d.f[0] * cos(2.0 * M_PI * d.f[0] * t) + d.f[1] * cos(2.0 * M_PI * d.f[1] * t) +
                d.f[2] * cos(2.0 * M_PI * d.f[2] * t));
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + d.a[k] * 2.0 * M_PI * d.f[k] * cos(2.0 * M_PI * d.f[k] * t + d.ph[k]);
    return s;
}

// (R', R'', tau', q') with every expression as the reference writes it.  A
// value the reference computes more than once from the same operands (p,
// p_in, I, tau', q' inside and outside F_R) is computed once here: the same
// operations on the same operands give the same bits.  pow(x, 2) is x*x,
// which equals a correctly rounded pow.
__device__ __forceinline__ void gm_rhs_ref(const GmRefModel& m, const GmRefDrive& d, double t, double R0,
                                           const double y[4], double f[4]) {
    const double R = y[0], R1 = y[1], Tau = y[2], q = y[3];
    const double p = ref_P(d, t);
    const double p_dt = ref_P_dt(d, t);
    const double p_in = (m.P0 + (m.S2 / R0)) * pow(R0 / R, m.gamma3);                        // G:82-85
    const double el = -m.G4_3 * (1.0 - pow(R0 / R, 3.0));
    const double Tau1 = (el - (m.mu4 * R1 / R) - Tau) / m.lambda;                              // G:106-110
    const double q1 = (el - (m.mu4 * R1 / R)) / (3.0 * m.lambda) - (q / m.lambda) - (R1 * Tau / R);  // G:113-117
    const double u = p_in - (m.S2 / R) + Tau + m.B;                                           // G:100, 125, 146
    const double v = m.P0 + p + m.B;
    const double I = u / v;
    const double C = m.C_inf * pow(I, (m.n - 1.0) / (2.0 * m.n));                              // G:96-103
    const double part1 = (m.P0 + p + m.B) * m.n / (m.rho * (m.n - 1.0));
    const double part2 = pow(I, (m.n - 1.0) / m.n) - 1.0;
    const double H = part1 * part2;                                                            // G:121-128
    const double part1_dt = p_dt * m.n / (m.rho * (m.n - 1.0));
    const double d_pin = (m.P0 + (m.S2 / R0)) * pow(R0, m.gamma3) * (-m.gamma3) * pow(R, -m.gamma3 - 1.0) * R1;  // G:89-92
    const double u_dt = d_pin + (m.S2 * R1 / (R * R)) + Tau1;
    const double I_dt = (u_dt * v - u * p_dt) / (v * v);
    const double part2_dt = ((m.n - 1.0) / m.n) * pow(I, -1.0 / m.n) * I_dt;
    const double H1 = part1_dt * part2 + part1 * part2_dt;                                    // G:132-157
    const double a = 1.0 - (R1 / C);
    const double b = 1.0 + (R1 / C);
    const double t1 = 1.5 * (R1 * R1) * (1.0 - (R1 / (3.0 * C)));
    const double t2 = b * (H + ((-Tau + 3.0 * q) / m.rho));
    const double t3 = (H1 * a + ((-Tau1 + 3.0 * q1) / m.rho)) * R / C;
    f[0] = R1;
    f[1] = (-t1 + t2 + t3) / (a * R);                                                          // G:161-179
    f[2] = Tau1;
    f[3] = q1;
}

// One DP5(4) attempt from (t, y) with step h and first stage k1 (G:246-294);
// rhs(t, y, f) evaluates the right-hand side.  Produces the 5th-order
// solution Y, the last stage k7 = f(t+h, .) and the RMS error norm.  Sums
// follow the reference's accumulation order.
template <typename F>
__device__ __forceinline__ double dp_attempt(const GmConsts& K, const F& rhs, double t, const double y[4], double h,
                                             const double k1[4], double Y[4], double k7[4]) {
    double k[7][4];
    double E[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        k[0][c] = k1[c];
        Y[c] = y[c] + K.tab.b[0] * k[0][c] * h;
        E[c] = k[0][c] * h * K.tab.e[0];
    }
#pragma unroll
    for (int i = 1; i < 7; ++i) {
        double ys[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < i; ++j) acc = acc + K.tab.a[i][j] * k[j][c] * h;
            ys[c] = y[c] + acc;
        }
        rhs(t + h * K.tab.c[i], ys, k[i]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            Y[c] = Y[c] + K.tab.b[i] * k[i][c] * h;
            E[c] = E[c] + k[i][c] * h * K.tab.e[i];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) k7[c] = k[6][c];
    double S = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double sc = K.abs_eps + K.rel_eps * fmax(fabs(y[c]), fabs(Y[c]));
        const double r = E[c] / sc;
        S = S + r * r;
    }
    return sqrt(S / 4.0);                                                   // G:183-192
}

// Classical RK4 step (same operation order as the oracle).
template <typename F>
__device__ __forceinline__ void rk4_step(const F& rhs, double t, double h, double y[4]) {
    double k1[4], k2[4], k3[4], k4[4], s[4];
    rhs(t, y, k1);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = y[i] + 0.5 * h * k1[i];
    rhs(t + 0.5 * h, s, k2);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = y[i] + 0.5 * h * k2[i];
    rhs(t + 0.5 * h, s, k3);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = y[i] + h * k3[i];
    rhs(t + h, s, k4);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = y[i] + h / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

// G:296: max(0.1, 0.9 err^-0.2)
__device__ __forceinline__ double step_factor(double err) { return fmax(0.1, 0.9 * pow(err, -0.2)); }

// ---------------------------------------------------------------------------
// batch integrator (persistent, per-lane refill)
// ---------------------------------------------------------------------------
// Integrator state of a bubble suspended at the end of a time slice.
struct GmLane {
    double y[4], k1[4];
    double t, h, rmax, tc, fac_max;
    long long acc, rej, s_rk;
    int j_out, flags;           // flags: 1 = k1 valid, 2 = at the start of a step
};

struct GmBatchArgs {
    GmConsts K;
    int method, precision, n, out_every, n_out, chunk;
    double w[3], f[3];
    double t_end, dt, split_ratio;
    long long max_steps;
    const double* R0; const double* amp; const double* phase;
    double* R_out; double* Rmax; double* t_collapse;
    long long* n_acc; long long* n_rej;
    int* counter;
    const int* order;        // queue position -> bubble index, nullable (input order)
    // time slicing: a bubble still running after `slice` attempts in this
    // launch is suspended into state[i], with its predicted remaining
    // attempts in cost[i] (-1 once it is finished); resume=1 continues
    // bubbles from state[i]
    GmLane* state;
    long long* cost;
    long long slice;
    int resume;
};

// MINW: wavefronts per SIMD the register allocation must allow (the
// DP5(4) kernels in the reference arithmetic and in the split take more than
// 256 registers and run one wavefront per SIMD; asked for two, they spill to
// scratch: faster while every lane has work, slower in the tail).
template <int METHOD, int PREC, int ARITH, int MINW = 1>
__global__ void __launch_bounds__(256, MINW) gm_batch_kernel(GmBatchArgs a) {
    static_assert(ARITH == BD_GM_ARITH_FAST || PREC == BD_GM_FP64, "reference arithmetic is FP64 only");
    const GmConsts& K = a.K;
    int next = 0, end = 0;          // this lane's current chunk of bubble indices
    bool have = false;
    int i = 0;
    GmDrive d;
    GmRefDrive rd;
    double R0 = 0, Pin0 = 0, t = 0, h = 0, rmax = 0, tc = -1, fac_max = 5.0;
    double y[4] = {0, 0, 0, 0}, k1[4] = {0, 0, 0, 0};
    bool k1_ok = false, at_start = true;
    long long acc = 0, rej = 0, s_rk = 0, nsteps_rk = 0;
    long long sl = 0;               // attempts of the current bubble in this launch
    int j_out = 0;
    auto f64 = [&](double tt, const double* s, double* o) {
        if constexpr (ARITH == BD_GM_ARITH_REFERENCE) gm_rhs_ref(K.ref, rd, tt, R0, s, o);
        else rhs_at<double>(K, d, tt, R0, Pin0, s, o);
    };
    auto f32 = [&](double tt, const double* s, double* o) { rhs_at<float>(K, d, tt, R0, Pin0, s, o); };
    while (true) {
        if (!have) {
            if (next >= end) {
                next = atomicAdd(a.counter, a.chunk);
                end = min(next + a.chunk, a.n);
                if (next >= a.n) break;
            }
            const int qpos = next++;
            i = a.order ? a.order[qpos] : qpos;
            have = true;
            sl = 0;
            R0 = a.R0[i];
            Pin0 = K.m64.P0 + K.m64.S2 / R0;
#pragma unroll
            for (int k = 0; k < 3; ++k) { d.w[k] = a.w[k]; d.A[k] = a.amp[3 * i + k]; d.ph[k] = a.phase[3 * i + k]; }
            rd.mode = 3;
            rd.A = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) { rd.f[k] = a.f[k]; rd.a[k] = d.A[k]; rd.ph[k] = d.ph[k]; }
            nsteps_rk = llrint(a.t_end / a.dt);
            if (a.resume) {
                const GmLane& L = a.state[i];
#pragma unroll
                for (int c = 0; c < 4; ++c) { y[c] = L.y[c]; k1[c] = L.k1[c]; }
                t = L.t; h = L.h; rmax = L.rmax; tc = L.tc; fac_max = L.fac_max;
                acc = L.acc; rej = L.rej; s_rk = L.s_rk; j_out = L.j_out;
                k1_ok = (L.flags & 1) != 0; at_start = (L.flags & 2) != 0;
            } else {
                y[0] = R0; y[1] = y[2] = y[3] = 0.0;
                t = 0.0; h = a.dt; rmax = R0; tc = -1.0; fac_max = 5.0;
                k1_ok = false; at_start = true;
                acc = rej = 0; s_rk = 0;
                j_out = 0;
                if (a.R_out && a.n_out > 0) a.R_out[(size_t)i * a.n_out + j_out++] = R0;
            }
        }
        bool done = false;
        // BD_GM_SPLIT: FP32 right-hand sides while every active lane of the
        // wavefront has R/R0 >= split_ratio (one precision per wavefront, so
        // the two paths are never both executed)
        bool use32 = PREC == BD_GM_SPLIT && y[0] >= a.split_ratio * R0;
        if constexpr (PREC == BD_GM_SPLIT) use32 = __all(use32) != 0;
        if constexpr (METHOD == BD_GM_RK4) {
            if (PREC == BD_GM_SPLIT && use32) rk4_step(f32, s_rk * a.dt, a.dt, y);
            else rk4_step(f64, s_rk * a.dt, a.dt, y);
            ++s_rk; ++acc;
            t = s_rk * a.dt;
            if (y[0] > rmax) rmax = y[0];
            if (a.R_out && a.out_every > 0 && s_rk % a.out_every == 0 && j_out < a.n_out)
                a.R_out[(size_t)i * a.n_out + j_out++] = y[0];
            if (y[1] <= -K.C_inf) { tc = t; done = true; }
            if (s_rk >= nsteps_rk || (a.max_steps > 0 && acc >= a.max_steps)) done = true;
        } else {
            if (!k1_ok) {
                if (PREC == BD_GM_SPLIT && use32) f32(t, y, k1);
                else f64(t, y, k1);
                k1_ok = true;
            }
            if (at_start) {
                if (t + h > a.t_end) h = a.t_end - t;
                fac_max = 5.0;
                at_start = false;
            }
            double Y[4], k7[4];
            double err;
            if (PREC == BD_GM_SPLIT && use32) err = dp_attempt(K, f32, t, y, h, k1, Y, k7);
            else err = dp_attempt(K, f64, t, y, h, k1, Y, k7);
            const double mr = step_factor(err);
            bool accept = !(err > 1.0);
            if (!accept) {
                fac_max = 0.9;
                h = h * fmin(fac_max, mr);
                ++rej;
                // h underflow or non-finite error: take the attempt as it is
                if (!(h > 0.0) || !isfinite(err)) accept = true;
            }
            if (accept) {
                t = t + h;
#pragma unroll
                for (int c = 0; c < 4; ++c) { y[c] = Y[c]; k1[c] = k7[c]; }
                // the reference recomputes the first stage at the new point (G:259-273)
                if (ARITH == BD_GM_ARITH_REFERENCE) k1_ok = false;
                h = h * fmin(fac_max, mr);
                at_start = true;
                ++acc;
                if (y[0] > rmax) rmax = y[0];
                if (y[1] <= -K.C_inf) { tc = t; done = true; }
                if (!(t < a.t_end) || !(h > 0.0) || !isfinite(y[0])) done = true;
                if (a.max_steps > 0 && acc + rej >= a.max_steps) done = true;
            }
        }
        ++sl;
        if (done) {
            if (a.R_out)
                for (; j_out < a.n_out; ++j_out) a.R_out[(size_t)i * a.n_out + j_out] = NAN;
            a.Rmax[i] = rmax;
            a.t_collapse[i] = tc;
            if (a.n_acc) a.n_acc[i] = acc;
            if (a.n_rej) a.n_rej[i] = rej;
            if (a.cost) a.cost[i] = -1;
            have = false;
        } else if (a.slice > 0 && sl >= a.slice) {
            GmLane& L = a.state[i];
#pragma unroll
            for (int c = 0; c < 4; ++c) { L.y[c] = y[c]; L.k1[c] = k1[c]; }
            L.t = t; L.h = h; L.rmax = rmax; L.tc = tc; L.fac_max = fac_max;
            L.acc = acc; L.rej = rej; L.s_rk = s_rk; L.j_out = j_out;
            L.flags = (k1_ok ? 1 : 0) | (at_start ? 2 : 0);
            // remaining attempts, extrapolated from the rate so far
            const double frac = t > 0.0 ? (a.t_end - t) / t : 1e6;
            a.cost[i] = (long long)fmin(4.0e18, fmax(1.0, (double)(acc + rej) * frac));
            have = false;
        }
    }
}

// ---------------------------------------------------------------------------
// drop-in threshold sweep (G:323-429), persistent with per-lane refill
// ---------------------------------------------------------------------------
// Search state of a (combo, R0) cell suspended at the end of a launch.
struct GmCell {
    double y[4], k1[4];
    double t, h, fac_max, A5, A4, A3;
    long long steps;
    int idx, stage, flags, pad;
};
constexpr int kCellK1 = 1, kCellStart = 2, kCellThr5 = 4, kCellThr4 = 8, kCellThr3 = 16;

struct GmSweepArgs {
    GmConsts K;
    int triple, N_fr_combo, N_R0;
    double minR0, stepR0;
    double* results_A;
    const double* Fr_combo;
    // device counters: [0] cells claimed, [1] cells completed, [2] suspended
    // cells taken up, [3 + b] cells suspended into buffer b (b = 0, 1)
    int* counter;
    int cell0, cell_end;       // cells idx in [cell0, cell_end), claimed in increasing order
    int* host_flags;           // nullable, host-mapped: [0] claimed, [1] completed, [2] stop request
    const GmCell* susp_in;     // cells to continue first (counter[3 + b_in] of them), nullable
    GmCell* susp_out;          // cells this launch suspends (buffer 1 - b_in), nullable if budget = 0
    int b_in;
    long long budget;          // DP5(4) attempts per lane in this launch, 0 = no limit
};

// System-scope access to the host-mapped progress words (vector memory ops).
__device__ __forceinline__ void host_store(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int host_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Reference drive (G:40-57) as amplitudes per tone; the dual mode uses fr1 and fr3.
__device__ __forceinline__ void ref_drive(GmDrive& d, bool triple, const double fr[3], double A) {
    for (int k = 0; k < 3; ++k) { d.w[k] = 2.0 * M_PI * fr[k]; d.A[k] = 0.0; d.ph[k] = 0.0; }
    if (!triple || (fr[0] == fr[1] && fr[0] == fr[2])) {
        d.A[0] = A;
    } else if (fr[0] == fr[1] || fr[1] == fr[2]) {
        d.A[0] = A / sqrt(2.0);
        d.A[2] = A / sqrt(2.0);
    } else {
        d.A[0] = d.A[1] = d.A[2] = A / sqrt(3.0);
    }
}

// The same drive in the reference's own form (G:40-78).
__device__ __forceinline__ void ref_drive_mode(GmRefDrive& d, bool triple, const double fr[3], double A) {
    d.A = A;
    for (int k = 0; k < 3; ++k) { d.f[k] = fr[k]; d.a[k] = 0.0; d.ph[k] = 0.0; }
    if (!triple || (fr[0] == fr[1] && fr[0] == fr[2])) d.mode = 0;
    else if (fr[0] == fr[1] || fr[1] == fr[2]) d.mode = 1;
    else d.mode = 2;
}

template <int ARITH>
__global__ void __launch_bounds__(64) gm_sweep_kernel(GmSweepArgs a) {
    const GmConsts& K = a.K;
    const double A_MAX = 1.0e7;
    bool have = false;
    int idx = 0;
    double fr[3] = {0, 0, 0};
    double R0 = 0, Pin0 = 0, endT = 0, t = 0, h = 0, fac_max = 5.0;
    double y[4] = {0, 0, 0, 0}, k1[4] = {0, 0, 0, 0};
    bool k1_ok = false, at_start = true;
    int stage = 0;          // 1, 2, 3: search stage of G:371-426
    double A5 = 0, A4 = 0, A3 = 0;
    bool thr5 = false, thr4 = true, thr3 = true;
    long long steps = 0;    // accepted steps of the current solve
    long long used = 0;     // attempts of this lane in this launch
    GmDrive d;
    GmRefDrive rd;
    auto rhs = [&](double tt, const double* s, double* o) {
        if constexpr (ARITH == BD_GM_ARITH_REFERENCE) gm_rhs_ref(K.ref, rd, tt, R0, s, o);
        else rhs_at<double>(K, d, tt, R0, Pin0, s, o);
    };
    auto set_amplitude = [&](double A) {
        if constexpr (ARITH == BD_GM_ARITH_REFERENCE) ref_drive_mode(rd, a.triple != 0, fr, A);
        else ref_drive(d, a.triple != 0, fr, A);
    };
    auto start_solve = [&](double A, bool reset_h) {
        if (reset_h) h = K.h0;
        t = 0.0; y[0] = R0; y[1] = y[2] = y[3] = 0.0;
        set_amplitude(A);
        k1_ok = false; at_start = true; steps = 0;
    };
    auto set_cell = [&](int id) {        // the cell's fixed inputs (G:332-353)
        const int ic = id / a.N_R0, ir = id % a.N_R0;
        R0 = a.minR0 + ir * a.stepR0;
        Pin0 = K.m64.P0 + K.m64.S2 / R0;
        fr[0] = a.Fr_combo[3 * ic]; fr[1] = a.Fr_combo[3 * ic + 1]; fr[2] = a.Fr_combo[3 * ic + 2];
        if (a.triple) {
            if (fr[0] <= fr[1] && fr[0] <= fr[2]) endT = 5.0 / fr[0];
            if (fr[1] <= fr[0] && fr[1] <= fr[2]) endT = 5.0 / fr[1];
            if (fr[2] <= fr[0] && fr[2] <= fr[1]) endT = 5.0 / fr[2];
        } else {
            endT = 5.0 / fr[0];
        }
    };
    auto finish_cell = [&](double result) {
        a.results_A[idx] = result;
        have = false;
        const int done = atomicAdd(a.counter + 1, 1) + 1;
        if (a.host_flags && (done & 63) == 0) host_store(a.host_flags + 1, done);
    };

    while (true) {
        if (!have) {
            if (a.budget > 0 && used >= a.budget) break;
            // a stop request ends the claiming
            if (a.host_flags && host_load(a.host_flags + 2) != 0) break;
            if (a.susp_in) {
                // continue the cells an earlier launch suspended first
                const int n_in = a.counter[3 + a.b_in];
                if (__hip_atomic_load(a.counter + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n_in) {
                    const int r = atomicAdd(a.counter + 2, 1);
                    if (r < n_in) {
                        const GmCell& c = a.susp_in[r];
                        idx = c.idx;
                        set_cell(idx);
#pragma unroll
                        for (int k = 0; k < 4; ++k) { y[k] = c.y[k]; k1[k] = c.k1[k]; }
                        t = c.t; h = c.h; fac_max = c.fac_max; A5 = c.A5; A4 = c.A4; A3 = c.A3;
                        steps = c.steps; stage = c.stage;
                        k1_ok = (c.flags & kCellK1) != 0; at_start = (c.flags & kCellStart) != 0;
                        thr5 = (c.flags & kCellThr5) != 0; thr4 = (c.flags & kCellThr4) != 0;
                        thr3 = (c.flags & kCellThr3) != 0;
                        set_amplitude(stage == 1 ? A5 : (stage == 2 ? A4 : A3));
                        have = true;
                    }
                }
            }
            if (!have) {
                idx = a.cell0 + atomicAdd(a.counter, 1);
                if (idx >= a.cell_end) break;
                if (a.host_flags && ((idx - a.cell0) & 255) == 0) host_store(a.host_flags, idx - a.cell0 + 1);
                have = true;
                set_cell(idx);
                h = K.h0;
                thr5 = false; thr4 = true; thr3 = true;
                A5 = A4 = A3 = 0.0;
                stage = 1;
                A5 = A5 + 1.0e5;
                start_solve(A5, false);
            }
        }
        // one DP5(4) attempt of the current solve
        if (!k1_ok) { rhs(t, y, k1); k1_ok = true; }
        if (at_start) {
            if (t + h > endT) h = endT - t;
            fac_max = 5.0;
            at_start = false;
        }
        double Y[4], k7[4];
        const double err = dp_attempt(K, rhs, t, y, h, k1, Y, k7);
        ++used;
        const double mr = step_factor(err);
        bool accept = !(err > 1.0);
        if (!accept) {
            fac_max = 0.9;
            h = h * fmin(fac_max, mr);
            if (!(h > 0.0) || !isfinite(err)) accept = true;
        }
        bool solve_over = false, thr = false;
        if (accept) {
            t = t + h;
#pragma unroll
            for (int c = 0; c < 4; ++c) { y[c] = Y[c]; k1[c] = k7[c]; }
            // the reference recomputes the first stage at the new point (G:259-273)
            if (ARITH == BD_GM_ARITH_REFERENCE) k1_ok = false;
            h = h * fmin(fac_max, mr);
            at_start = true;
            ++steps;
            thr = Y[1] <= -K.C_inf;
            solve_over = !(t < endT && !thr && steps < K.cap);
        }
        if (solve_over) {
            if (!thr && t < endT) {
                // the solve reached the step cap undecided: the search cannot go on
                finish_cell(BD_GM_SWEEP_CAPPED);
                continue;
            }
            // advance the search (G:372-426)
            bool next_solve = false;
            double result = -1.0;
            if (stage == 1) {
                thr5 = thr;
                if (!thr5 && A5 < A_MAX) { A5 = A5 + 1.0e5; start_solve(A5, false); next_solve = true; }
                else {
                    if (!thr5 && A5 == A_MAX) result = A5;
                    A4 = A5;
                    if (thr5 && thr4) { stage = 2; A4 = A4 - 1.0e4; thr4 = false; start_solve(A4, true); next_solve = true; }
                    else stage = 3;
                }
            } else if (stage == 2) {
                thr4 = thr;
                if (thr4) { A4 = A4 - 1.0e4; thr4 = false; start_solve(A4, true); next_solve = true; }
                else {
                    A4 = A4 + 1.0e4;
                    A3 = A4;
                    if (thr5 && !thr4 && thr3) { stage = 3; A3 = A3 - 1.0e3; thr3 = false; start_solve(A3, true); next_solve = true; }
                }
            } else {
                thr3 = thr;
                if (!thr3) result = A3 + 1.0e3;
                else { A3 = A3 - 1.0e3; thr3 = false; start_solve(A3, true); next_solve = true; }
            }
            if (!next_solve) { finish_cell(result); continue; }
        }
        // out of budget, or asked to stop: suspend the cell for a later launch
        const bool stop = a.host_flags && (used & 255) == 0 && host_load(a.host_flags + 2) != 0;
        if (a.susp_out && ((a.budget > 0 && used >= a.budget) || stop)) {
            GmCell c;
#pragma unroll
            for (int k = 0; k < 4; ++k) { c.y[k] = y[k]; c.k1[k] = k1[k]; }
            c.t = t; c.h = h; c.fac_max = fac_max; c.A5 = A5; c.A4 = A4; c.A3 = A3;
            c.steps = steps; c.idx = idx; c.stage = stage; c.pad = 0;
            c.flags = (k1_ok ? kCellK1 : 0) | (at_start ? kCellStart : 0) | (thr5 ? kCellThr5 : 0) |
                      (thr4 ? kCellThr4 : 0) | (thr3 ? kCellThr3 : 0);
            const int pos = atomicAdd(a.counter + 3 + (1 - a.b_in), 1);
            a.susp_out[pos] = c;
            have = false;
            if (stop) break;
        }
    }
}

// After a range launch: the exact completed count and the finished mark.
__global__ void gm_sweep_final_kernel(const int* counter, int* host_flags) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        host_store(host_flags + 1, counter[1]);
        host_store(host_flags + 3, 1);
    }
}

// After an advance of a resumable sweep: claimed (clamped to the range),
// completed, suspended by this launch, then the finished mark.
__global__ void gm_sweep_advance_final_kernel(const int* counter, int b_out, int range, int* host_flags) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        host_store(host_flags + 0, min(counter[0], range));
        host_store(host_flags + 1, counter[1]);
        host_store(host_flags + 4, counter[3 + b_out]);
        host_store(host_flags + 3, 1);
    }
}

// Before an advance: the resume cursor and the output buffer's count start at 0.
__global__ void gm_sweep_advance_init_kernel(int* counter, int b_out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counter[2] = 0;
        counter[3 + b_out] = 0;
    }
}

// zeroes p[0 .. blockDim.x) (a fill kernel: safe inside captured graphs)
__global__ void zero_int_kernel(int* p) { if (blockIdx.x == 0) p[threadIdx.x] = 0; }

static GmConsts make_consts(const bd_gilmore_params* g) {
    GmConsts K;
    K.m64 = make_model<double>(g);
    K.m32 = make_model<float>(g);
    K.ref = make_ref_model(g);
    K.tab = make_tab(g->tableau);
    K.C_inf = g->C_inf;
    K.abs_eps = g->abs_eps;
    K.rel_eps = g->rel_eps;
    K.h0 = g->h0;
    K.cap = g->max_solve_steps > 0 ? (long long)g->max_solve_steps : 10000000LL;
    return K;
}

// Persistent grid: `waves` wavefronts per SIMD (4 SIMDs per CU) on every CU.
static int persistent_blocks(int block, int waves = 8) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) == hipSuccess && pr.multiProcessorCount > 0) cus = pr.multiProcessorCount;
    }
    return cus * std::max(1, waves * 4 * kWave / block);
}

// Launch shape of the batch integrator: workgroup size, bubbles per queue
// grab, wavefronts per SIMD of the persistent grid.  bd_gilmore_autotune
// measures the candidates and records the fastest per device, method and
// precision; bd_gilmore_integrate uses it for fields left at 0.
struct GmShape { int block, chunk, waves; };
constexpr GmShape kGmDefaultShape = {256, 1, 8};
constexpr int kGmMaxDevices = 64;
// indexed [device][method][precision][arithmetic]
static GmShape g_gm_tuned[kGmMaxDevices][2][2][2];
static bool g_gm_has_tuned[kGmMaxDevices][2][2][2];
// guards the per-device host-side state of this file (tuned shapes, scratch
// and counter allocations) against concurrent host threads
static std::mutex g_gm_mutex;

static GmShape gm_shape(const bd_gilmore_params* prm, const bd_gilmore_batch* b) {
    GmShape s = kGmDefaultShape;
    int dev = 0;
    std::lock_guard<std::mutex> lock(g_gm_mutex);
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kGmMaxDevices &&
        g_gm_has_tuned[dev][b->method][b->precision][prm->arithmetic])
        s = g_gm_tuned[dev][b->method][b->precision][prm->arithmetic];
    if (b->block > 0) s.block = b->block;
    if (b->chunk > 0) s.chunk = b->chunk;
    if (b->waves > 0) s.waves = b->waves;
    return s;
}

// the drop-in sweep's counters, one set per device (its signature has no workspace)
constexpr int kSweepCounters = 5;
// cells of one sweep: claims are counted in 32-bit words, and every lane of
// the persistent grid (2^19 at most) may claim once past the end
constexpr long long kSweepMaxCells = 0x7fffffffLL - (1LL << 24);
static int* sweep_counter() {
    static int* ctr[64] = {nullptr};
    std::lock_guard<std::mutex> lock(g_gm_mutex);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!ctr[dev] && hipMalloc(&ctr[dev], kSweepCounters * sizeof(int)) != hipSuccess) ctr[dev] = nullptr;
    return ctr[dev];
}

// Host-mapped progress words of the range sweep, one set per process:
// host pointer for the caller, device pointer for the kernel.
static int* g_sweep_flags_host = nullptr;
static int* g_sweep_flags_dev = nullptr;
static bool sweep_flags(int** host, int** dev) {
    std::lock_guard<std::mutex> lock(g_gm_mutex);
    if (!g_sweep_flags_host) {
        void* p = nullptr;
        if (hipHostMalloc(&p, 4 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return false;
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) { (void)hipHostFree(p); return false; }
        g_sweep_flags_host = (int*)p;
        g_sweep_flags_dev = (int*)d;
        std::memset(p, 0, 4 * sizeof(int));
    }
    *host = g_sweep_flags_host;
    *dev = g_sweep_flags_dev;
    return true;
}

}  // namespace bd

using namespace bd;

extern "C" void bd_gilmore_reference_params(bd_gilmore_params* p) {  // G:440-460
    if (!p) return;
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

static int check_gilmore(const bd_gilmore_params* g) {
    if (!g) return BD_ERR_INVALID;
    if (!(g->rho > 0) || !(g->lambda > 0) || !(g->n > 1) || !(g->C_inf > 0) || !(g->h0 > 0))
        return BD_ERR_INVALID;
    if (g->tableau != BD_GM_TAB_REFERENCE && g->tableau != BD_GM_TAB_DP1980) return BD_ERR_INVALID;
    if (g->arithmetic != BD_GM_ARITH_REFERENCE && g->arithmetic != BD_GM_ARITH_FAST) return BD_ERR_INVALID;
    return BD_OK;
}

// Launch geometry of the sweep kernel: 64-lane workgroups, 8 waves per SIMD.
constexpr int kSweepBlock = 64;
static int sweep_blocks(long long cells) {
    const int blocks = persistent_blocks(kSweepBlock);
    return (int)std::max<long long>(1, std::min<long long>(blocks, (cells + kSweepBlock - 1) / kSweepBlock));
}

static void sweep_args(GmSweepArgs& a, const bd_gilmore_params* prm, double* results_A, const double* Fr_combo,
                       int N_fr_combo, int N_R0, int* ctr, int cell0, int cell_end) {
    a.K = make_consts(prm);
    a.triple = prm->triple_fr;
    a.N_fr_combo = N_fr_combo;
    a.N_R0 = N_R0;
    a.minR0 = prm->minR0;
    a.stepR0 = prm->stepR0;
    a.results_A = results_A;
    a.Fr_combo = Fr_combo;
    a.counter = ctr;
    a.cell0 = cell0;
    a.cell_end = cell_end;
    a.host_flags = nullptr;
    a.susp_in = nullptr;
    a.susp_out = nullptr;
    a.b_in = 0;
    a.budget = 0;
}

static int launch_sweep_kernel(const GmSweepArgs& a, int arith, int blocks, hipStream_t st) {
    if (arith == BD_GM_ARITH_REFERENCE)
        hipLaunchKernelGGL(gm_sweep_kernel<BD_GM_ARITH_REFERENCE>, dim3(blocks), dim3(kSweepBlock), 0, st, a);
    else
        hipLaunchKernelGGL(gm_sweep_kernel<BD_GM_ARITH_FAST>, dim3(blocks), dim3(kSweepBlock), 0, st, a);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

// One unbudgeted launch over [cell0, cell_end) from zeroed counters.
static int launch_sweep(GmSweepArgs& a, int arith, hipStream_t st) {
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(kSweepCounters), 0, st, a.counter);
    BD_LAUNCH_CHECK();
    return launch_sweep_kernel(a, arith, sweep_blocks((long long)a.cell_end - a.cell0), st);
}

static int check_sweep(const bd_gilmore_params* prm, const double* results_A, const double* Fr_combo,
                       int N_fr_combo, int N_R0) {
    int rc = check_gilmore(prm);
    if (rc) return rc;
    if (!results_A || !Fr_combo || N_fr_combo <= 0 || N_R0 <= 0) return BD_ERR_INVALID;
    if ((long long)N_fr_combo * N_R0 > kSweepMaxCells) return BD_ERR_INVALID;
    return BD_OK;
}

extern "C" int bd_gilmore_threshold_sweep(const bd_gilmore_params* prm, double* results_A,
                                          const double* Fr_combo, int N_fr_combo, int N_R0, int N_A,
                                          void* stream) {
    (void)N_A;
    int rc = check_sweep(prm, results_A, Fr_combo, N_fr_combo, N_R0);
    if (rc) return rc;
    int* ctr = sweep_counter();
    if (!ctr) return BD_ERR_ALLOC;
    GmSweepArgs a;
    sweep_args(a, prm, results_A, Fr_combo, N_fr_combo, N_R0, ctr, 0, N_fr_combo * N_R0);
    return launch_sweep(a, prm->arithmetic, as_stream(stream));
}

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; nullptr if the lookup fails.
static int* noalloc_scratch_ptr() {
    static int* ptr[kGmMaxDevices] = {nullptr};
    std::lock_guard<std::mutex> lock(g_gm_mutex);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kGmMaxDevices) return nullptr;
    if (!ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            fprintf(stderr, "bubbledynamics: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return nullptr;
        }
        ptr[dev] = (int*)p;
    }
    return ptr[dev];
}

extern "C" int bd_gilmore_threshold_sweep_noalloc(const bd_gilmore_params* prm, double* results_A,
                                                  const double* Fr_combo, int N_fr_combo, int N_R0, int N_A,
                                                  void* stream) {
    (void)N_A;
    int rc = check_gilmore(prm);
    if (rc) return rc;
    if (!results_A || !Fr_combo || N_fr_combo <= 0 || N_R0 <= 0) return BD_ERR_INVALID;
    if ((long long)N_fr_combo * N_R0 > kSweepMaxCells) {
        fprintf(stderr, "bubbledynamics: bd_gilmore_threshold_sweep_noalloc: %lld cells exceed the maximum %lld "
                        "(32-bit queue counters)\n", (long long)N_fr_combo * N_R0, (long long)kSweepMaxCells);
        return BD_ERR_INVALID;
    }
    int* ctr = noalloc_scratch_ptr();
    if (!ctr) return BD_ERR_HIP;
    GmSweepArgs a;
    sweep_args(a, prm, results_A, Fr_combo, N_fr_combo, N_R0, ctr, 0, N_fr_combo * N_R0);
    // launch_sweep clears the counters with a fill kernel before the sweep kernel
    return launch_sweep(a, prm->arithmetic, as_stream(stream));
}

extern "C" int bd_gilmore_threshold_sweep_range(const bd_gilmore_params* prm, double* results_A,
                                                const double* Fr_combo, int N_fr_combo, int N_R0, int cell0,
                                                int cell_end, void* stream) {
    int rc = check_sweep(prm, results_A, Fr_combo, N_fr_combo, N_R0);
    if (rc) return rc;
    if (cell0 < 0 || cell_end > N_fr_combo * N_R0 || cell0 >= cell_end) return BD_ERR_INVALID;
    int* ctr = sweep_counter();
    if (!ctr) return BD_ERR_ALLOC;
    int *hf = nullptr, *df = nullptr;
    if (!sweep_flags(&hf, &df)) return BD_ERR_ALLOC;
    hipStream_t st = as_stream(stream);
    // the caller's stop request (hf[2]) is left as it is: a stop set before the
    // call makes the launch claim nothing
    hf[0] = 0; hf[1] = 0; hf[3] = 0;
    GmSweepArgs a;
    sweep_args(a, prm, results_A, Fr_combo, N_fr_combo, N_R0, ctr, cell0, cell_end);
    a.host_flags = df;
    rc = launch_sweep(a, prm->arithmetic, st);
    if (rc) return rc;
    hipLaunchKernelGGL(gm_sweep_final_kernel, dim3(1), dim3(1), 0, st, (const int*)ctr, df);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

// Resumable sweep: the drop-in's kernel in launches of bounded length.
struct bd_gilmore_sweep {
    bd::GmSweepArgs base;
    int arith = 0;
    int device = 0;
    int blocks = 0;            // persistent grid of every advance (suspension capacity = lanes)
    int* d_ctr = nullptr;      // kSweepCounters words
    bd::GmCell* susp[2] = {nullptr, nullptr};
    int b_in = 0;              // buffer holding the cells the next advance continues
    int* h_flags = nullptr;    // host-mapped: [0] claimed [1] completed [2] stop [3] launch ended [4] suspended
    int* d_flags = nullptr;
};

extern "C" int bd_gilmore_sweep_destroy(bd_gilmore_sweep* sw) {
    if (!sw) return BD_OK;
    (void)hipDeviceSynchronize();
    if (sw->d_ctr) (void)hipFree(sw->d_ctr);
    for (int b = 0; b < 2; ++b) if (sw->susp[b]) (void)hipFree(sw->susp[b]);
    if (sw->h_flags) (void)hipHostFree(sw->h_flags);
    delete sw;
    return BD_OK;
}

extern "C" int bd_gilmore_sweep_create(const bd_gilmore_params* prm, const double* Fr_combo, int N_fr_combo,
                                       int N_R0, int cell0, int cell_end, bd_gilmore_sweep** out) {
    using namespace bd;
    if (!out) return BD_ERR_INVALID;
    *out = nullptr;
    int rc = check_gilmore(prm);
    if (rc) return rc;
    if (!Fr_combo || N_fr_combo <= 0 || N_R0 <= 0 || (long long)N_fr_combo * N_R0 > kSweepMaxCells)
        return BD_ERR_INVALID;
    if (cell0 < 0 || cell_end > N_fr_combo * N_R0 || cell0 >= cell_end) return BD_ERR_INVALID;
    auto* sw = new bd_gilmore_sweep();
    sw->arith = prm->arithmetic;
    sweep_args(sw->base, prm, nullptr, Fr_combo, N_fr_combo, N_R0, nullptr, cell0, cell_end);
    sw->blocks = sweep_blocks((long long)cell_end - cell0);
    const size_t cap = (size_t)sw->blocks * kSweepBlock;
    void* hp = nullptr;
    void* dp = nullptr;
    if (hipGetDevice(&sw->device) != hipSuccess ||
        hipMalloc(&sw->d_ctr, kSweepCounters * sizeof(int)) != hipSuccess ||
        hipMalloc(&sw->susp[0], cap * sizeof(GmCell)) != hipSuccess ||
        hipMalloc(&sw->susp[1], cap * sizeof(GmCell)) != hipSuccess ||
        hipHostMalloc(&hp, 8 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        bd_gilmore_sweep_destroy(sw);
        return BD_ERR_ALLOC;
    }
    sw->h_flags = (int*)hp;
    std::memset(hp, 0, 8 * sizeof(int));
    if (hipHostGetDevicePointer(&dp, hp, 0) != hipSuccess ||
        hipMemset(sw->d_ctr, 0, kSweepCounters * sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        bd_gilmore_sweep_destroy(sw);
        return BD_ERR_HIP;
    }
    sw->d_flags = (int*)dp;
    sw->h_flags[3] = 1;
    *out = sw;
    return BD_OK;
}

extern "C" int bd_gilmore_sweep_advance(bd_gilmore_sweep* sw, double* results_A, long long budget, void* stream) {
    using namespace bd;
    if (!sw || !results_A || budget < 0) return BD_ERR_INVALID;
    int dev = 0;
    BD_HIP_CHECK(hipGetDevice(&dev));
    if (dev != sw->device) return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    const int b_out = 1 - sw->b_in;
    GmSweepArgs a = sw->base;
    a.results_A = results_A;
    a.counter = sw->d_ctr;
    a.host_flags = sw->d_flags;
    a.susp_in = sw->susp[sw->b_in];
    a.susp_out = sw->susp[b_out];
    a.b_in = sw->b_in;
    a.budget = budget;
    sw->h_flags[3] = 0;
    hipLaunchKernelGGL(gm_sweep_advance_init_kernel, dim3(1), dim3(1), 0, st, sw->d_ctr, b_out);
    BD_LAUNCH_CHECK();
    int rc = launch_sweep_kernel(a, sw->arith, sw->blocks, st);
    if (rc) return rc;
    hipLaunchKernelGGL(gm_sweep_advance_final_kernel, dim3(1), dim3(1), 0, st, (const int*)sw->d_ctr, b_out,
                       a.cell_end - a.cell0, sw->d_flags);
    BD_LAUNCH_CHECK();
    sw->b_in = b_out;
    return BD_OK;
}

extern "C" int bd_gilmore_sweep_progress(const bd_gilmore_sweep* sw, long long out[5]) {
    if (!sw || !out) return BD_ERR_INVALID;
    const volatile int* f = sw->h_flags;
    const int range = sw->base.cell_end - sw->base.cell0;
    out[0] = f[0];
    out[1] = f[1];
    out[2] = f[4];
    out[3] = f[1] >= range ? 1 : 0;
    out[4] = f[3];
    return BD_OK;
}

extern "C" int* bd_gilmore_sweep_stop_word(bd_gilmore_sweep* sw) { return sw ? sw->h_flags + 2 : nullptr; }

extern "C" int* bd_gilmore_sweep_flags(void) {
    int *hf = nullptr, *df = nullptr;
    return sweep_flags(&hf, &df) ? hf : nullptr;
}

extern "C" size_t bd_gilmore_workspace_bytes(void) { return 256; }

// Queue order of the second pass: quarter-octave buckets of the predicted
// remaining attempts, most expensive bucket first, finished bubbles left
// out (order inside a bucket is the scatter's arrival order; the results do
// not depend on it).  hist[kGmBuckets] receives the number of bubbles queued.
constexpr int kGmBuckets = 128;
__device__ __forceinline__ int gm_bucket(long long cost) {
    const int k = (int)(4.0 * log2((double)cost + 1.0));
    return k < 0 ? 0 : (k >= kGmBuckets ? kGmBuckets - 1 : k);
}
__global__ void gm_zero_buckets(int* hist) {
    if (threadIdx.x <= kGmBuckets) hist[threadIdx.x] = 0;
}
__global__ void gm_bucket_hist(const long long* cost, int n, int* hist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cost[i] >= 0) atomicAdd(&hist[gm_bucket(cost[i])], 1);
}
__global__ void gm_bucket_starts(int* hist) {   // counts -> start positions, top bucket first
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int run = 0;
    for (int k = kGmBuckets - 1; k >= 0; --k) { const int c = hist[k]; hist[k] = run; run += c; }
    hist[kGmBuckets] = run;
}
__global__ void gm_bucket_scatter(const long long* cost, int n, int* cursor, int* order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cost[i] >= 0) order[atomicAdd(&cursor[gm_bucket(cost[i])], 1)] = i;
}

// Scratch of the sliced schedule, one growing buffer per device, held until
// the process ends: state[n], cost[n], order[n], bucket table.
struct GmSched { GmLane* state; long long* cost; int* order; int* hist; };
static bool gm_sched_scratch(int n, GmSched* out) {
    static void* buf[kGmMaxDevices] = {nullptr};
    static size_t cap[kGmMaxDevices] = {0};
    std::lock_guard<std::mutex> lock(g_gm_mutex);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kGmMaxDevices) return false;
    const size_t s_state = sizeof(GmLane) * (size_t)n, s_cost = sizeof(long long) * (size_t)n;
    const size_t s_order = sizeof(int) * (size_t)n, s_hist = sizeof(int) * (kGmBuckets + 1);
    const size_t need = s_state + s_cost + s_order + s_hist;
    if (cap[dev] < need) {
        if (buf[dev]) { (void)hipDeviceSynchronize(); (void)hipFree(buf[dev]); buf[dev] = nullptr; cap[dev] = 0; }
        if (hipMalloc(&buf[dev], need) != hipSuccess) return false;
        cap[dev] = need;
    }
    char* p = (char*)buf[dev];
    out->state = (GmLane*)p;
    out->cost = (long long*)(p + s_state);
    out->order = (int*)(p + s_state + s_cost);
    out->hist = (int*)(p + s_state + s_cost + s_order);
    return true;
}

struct GmPass {                  // order / slicing of one launch of the batch kernel
    const int* order = nullptr;
    GmLane* state = nullptr;
    long long* cost = nullptr;
    long long slice = 0;
    int resume = 0;
    // the DP5(4) kernels of the reference arithmetic and of the split in
    // their build for two wavefronts per SIMD (same results)
    int two_waves = 0;
};

static int gm_launch(const bd_gilmore_params* prm, const bd_gilmore_batch* b, void* workspace, GmShape shp,
                     hipStream_t st, const GmPass& pass = GmPass()) {
    if (shp.block % kWave != 0 || shp.block > 256 || shp.block <= 0 || shp.chunk <= 0 || shp.waves <= 0)
        return BD_ERR_INVALID;
    GmBatchArgs a;
    a.K = make_consts(prm);
    if (b->method == BD_GM_DP45) { a.K.abs_eps = b->atol; a.K.rel_eps = b->rtol; }
    a.method = b->method;
    a.precision = b->precision;
    a.n = b->n;
    a.out_every = b->out_every;
    a.n_out = b->R_out ? b->n_out : 0;
    a.chunk = shp.chunk;
    for (int k = 0; k < 3; ++k) { a.w[k] = 2.0 * M_PI * b->freq[k]; a.f[k] = b->freq[k]; }
    a.t_end = b->t_end;
    a.dt = b->dt;
    a.split_ratio = b->split_ratio;
    a.max_steps = b->max_steps > 0 ? b->max_steps : 10000000LL;   // every lane terminates
    a.R0 = b->R0; a.amp = b->amp; a.phase = b->phase;
    a.R_out = b->R_out; a.Rmax = b->Rmax; a.t_collapse = b->t_collapse;
    a.n_acc = b->n_acc; a.n_rej = b->n_rej;
    a.counter = (int*)workspace;
    a.order = pass.order;
    a.state = pass.state;
    a.cost = pass.cost;
    a.slice = pass.slice;
    a.resume = pass.resume;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, st, a.counter);
    BD_LAUNCH_CHECK();
    const int block = shp.block;
    int blocks = persistent_blocks(block, shp.waves);
    const long long need = ((long long)b->n + (long long)a.chunk - 1) / a.chunk;
    blocks = (int)std::min<long long>(blocks, (need + block - 1) / block);
    const dim3 grid(std::max(1, blocks)), blk(block);
    constexpr int REF = BD_GM_ARITH_REFERENCE, FAST = BD_GM_ARITH_FAST;
    const bool ref = prm->arithmetic == REF;
    if (b->method == BD_GM_RK4 && b->precision == BD_GM_FP64 && ref)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_RK4, BD_GM_FP64, REF>), grid, blk, 0, st, a);
    else if (b->method == BD_GM_RK4 && b->precision == BD_GM_FP64)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_RK4, BD_GM_FP64, FAST>), grid, blk, 0, st, a);
    else if (b->method == BD_GM_RK4)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_RK4, BD_GM_SPLIT, FAST>), grid, blk, 0, st, a);
    else if (b->precision == BD_GM_FP64 && ref && pass.two_waves)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_DP45, BD_GM_FP64, REF, 2>), grid, blk, 0, st, a);
    else if (b->precision == BD_GM_FP64 && ref)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_DP45, BD_GM_FP64, REF>), grid, blk, 0, st, a);
    else if (b->precision == BD_GM_FP64)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_DP45, BD_GM_FP64, FAST>), grid, blk, 0, st, a);
    else if (pass.two_waves)
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_DP45, BD_GM_SPLIT, FAST, 2>), grid, blk, 0, st, a);
    else
        hipLaunchKernelGGL((gm_batch_kernel<BD_GM_DP45, BD_GM_SPLIT, FAST>), grid, blk, 0, st, a);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

static int check_batch(const bd_gilmore_params* prm, const bd_gilmore_batch* b) {
    int rc = check_gilmore(prm);
    if (rc) return rc;
    if (!b || b->n <= 0 || !b->R0 || !b->amp || !b->phase || !b->Rmax || !b->t_collapse || !b->n_acc || !b->n_rej)
        return BD_ERR_INVALID;
    if (b->method != BD_GM_RK4 && b->method != BD_GM_DP45) return BD_ERR_INVALID;
    if (b->precision != BD_GM_FP64 && b->precision != BD_GM_SPLIT) return BD_ERR_INVALID;
    // the reference's arithmetic is FP64 throughout
    if (b->precision == BD_GM_SPLIT && prm->arithmetic == BD_GM_ARITH_REFERENCE) return BD_ERR_UNSUPPORTED;
    if (!(b->dt > 0) || !(b->t_end > 0)) return BD_ERR_INVALID;
    if (b->R_out && b->n_out <= 0) return BD_ERR_INVALID;
    if (b->method == BD_GM_DP45 && (!(b->rtol > 0) || !(b->atol >= 0))) return BD_ERR_INVALID;
    if (b->block < 0 || b->chunk < 0 || b->waves < 0) return BD_ERR_INVALID;
    return BD_OK;
}

extern "C" int bd_gilmore_integrate(const bd_gilmore_params* prm, const bd_gilmore_batch* b, void* workspace,
                                    void* stream) {
    int rc = check_batch(prm, b);
    if (rc) return rc;
    if (!workspace) return BD_ERR_INVALID;
    if (b->schedule != BD_GM_SCHED_INPUT && b->schedule != BD_GM_SCHED_SLICED) return BD_ERR_INVALID;
    if (b->slice < 0) return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    const GmShape shp = gm_shape(prm, b);
    if (b->schedule == BD_GM_SCHED_INPUT) return gm_launch(prm, b, workspace, shp, st);
    // Passes of growing slices (S, 16 S, then to the end): every pass runs the
    // bubbles still unfinished, largest predicted remaining work first, so a
    // long bubble the previous ranking underrated is moved up after one slice.
    // The two sliced passes keep every lane busy and take the kernels built
    // for two wavefronts per SIMD; the last pass ends in a tail of a few long
    // bubbles, where a wavefront's own speed counts, and takes the builds that
    // do not spill (1e6 bubbles, 20 us, reference arithmetic: passes of
    // 72 / 756 / 1,381 ms with one wavefront per SIMD, 59 / 633 / 1,565 ms
    // with two).
    GmSched sc;
    if (!gm_sched_scratch(b->n, &sc)) return BD_ERR_ALLOC;
    const long long S = b->slice > 0 ? b->slice : 128;
    const long long slices[3] = {S, 16 * S, 0};
    const unsigned g = (unsigned)((b->n + 255) / 256);
    int m = b->n;
    for (int pass = 0; pass < 3 && m > 0; ++pass) {
        bd_gilmore_batch bp = *b;
        bp.n = m;
        GmPass pp;
        pp.order = pass > 0 ? sc.order : nullptr;
        pp.state = sc.state;
        pp.cost = sc.cost;
        pp.slice = slices[pass];
        pp.resume = pass > 0;
        pp.two_waves = slices[pass] != 0;
        rc = gm_launch(prm, &bp, workspace, shp, st, pp);
        if (rc) return rc;
        if (slices[pass] == 0) break;
        // rank the unfinished bubbles (cost >= 0) by predicted remaining attempts
        hipLaunchKernelGGL(gm_zero_buckets, dim3(1), dim3(kGmBuckets + 1), 0, st, sc.hist);
        hipLaunchKernelGGL(gm_bucket_hist, dim3(g), dim3(256), 0, st, (const long long*)sc.cost, b->n, sc.hist);
        hipLaunchKernelGGL(gm_bucket_starts, dim3(1), dim3(1), 0, st, sc.hist);
        hipLaunchKernelGGL(gm_bucket_scatter, dim3(g), dim3(256), 0, st, (const long long*)sc.cost, b->n, sc.hist,
                           sc.order);
        BD_LAUNCH_CHECK();
        BD_HIP_CHECK(hipMemcpyAsync(&m, sc.hist + kGmBuckets, sizeof(int), hipMemcpyDeviceToHost, st));
        BD_HIP_CHECK(hipStreamSynchronize(st));
    }
    return BD_OK;
}

// Autotuner: times every candidate launch shape on the first `sample`
// bubbles of the batch (each capped at probe_steps attempts, so every lane
// has work for the whole run and the rate measures steady-state throughput),
// keeps the fastest for this device / method / precision and returns it.
// The integrator's results do not depend on the launch shape.
extern "C" int bd_gilmore_autotune(const bd_gilmore_params* prm, const bd_gilmore_batch* batch, int sample,
                                   long long probe_steps, void* stream, int32_t* best_block, int32_t* best_chunk,
                                   int32_t* best_waves, double* best_rate) {
    int rc = check_batch(prm, batch);
    if (rc) return rc;
    if (sample < 0 || probe_steps < 0) return BD_ERR_INVALID;
    int dev = 0;
    BD_HIP_CHECK(hipGetDevice(&dev));
    if (dev >= kGmMaxDevices) return BD_ERR_UNSUPPORTED;
    hipStream_t st = as_stream(stream);
    const int ns = std::min(batch->n, sample > 0 ? sample : 1 << 20);
    bd_gilmore_batch b = *batch;
    b.n = ns;
    b.R_out = nullptr; b.n_out = 0; b.out_every = 0;
    b.max_steps = probe_steps > 0 ? probe_steps : 500;
    struct Scratch {
        void* p = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Scratch() {
            if (p) (void)hipFree(p);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } sc;
    BD_HIP_CHECK(hipMalloc(&sc.p, (size_t)ns * 32 + 256));
    BD_HIP_CHECK(hipEventCreate(&sc.e0));
    BD_HIP_CHECK(hipEventCreate(&sc.e1));
    char* base = (char*)sc.p;
    b.Rmax = (double*)base; b.t_collapse = (double*)(base + (size_t)ns * 8);
    b.n_acc = (long long*)(base + (size_t)ns * 16); b.n_rej = (long long*)(base + (size_t)ns * 24);
    void* counter = base + (size_t)ns * 32;
    const GmShape cands[] = {
        {64, 1, 2}, {64, 1, 4}, {64, 1, 8}, {128, 1, 2}, {128, 1, 4}, {128, 1, 8},
        {256, 1, 2}, {256, 1, 4}, {256, 1, 8}, {64, 4, 8}, {128, 4, 8}, {256, 4, 8},
    };
    // warm-up with the default shape, and the step total (the same for every shape)
    rc = gm_launch(prm, &b, counter, kGmDefaultShape, st);
    if (rc) return rc;
    std::vector<long long> acc((size_t)ns);
    BD_HIP_CHECK(hipMemcpyAsync(acc.data(), b.n_acc, (size_t)ns * 8, hipMemcpyDeviceToHost, st));
    BD_HIP_CHECK(hipStreamSynchronize(st));
    double steps = 0.0;
    for (long long v : acc) steps += (double)v;
    GmShape best = kGmDefaultShape;
    double best_ms = 1e300;
    for (const GmShape& c : cands) {
        BD_HIP_CHECK(hipEventRecord(sc.e0, st));
        rc = gm_launch(prm, &b, counter, c, st);
        if (rc) return rc;
        BD_HIP_CHECK(hipEventRecord(sc.e1, st));
        BD_HIP_CHECK(hipEventSynchronize(sc.e1));
        float t = 0.f;
        BD_HIP_CHECK(hipEventElapsedTime(&t, sc.e0, sc.e1));
        if (t < best_ms) { best_ms = t; best = c; }
    }
    {
        std::lock_guard<std::mutex> lock(g_gm_mutex);
        g_gm_tuned[dev][b.method][b.precision][prm->arithmetic] = best;
        g_gm_has_tuned[dev][b.method][b.precision][prm->arithmetic] = true;
    }
    if (best_block) *best_block = best.block;
    if (best_chunk) *best_chunk = best.chunk;
    if (best_waves) *best_waves = best.waves;
    if (best_rate) *best_rate = steps / (best_ms * 1e-3);
    return BD_OK;
}

// Threshold table in the reference's output format (G:580-606): a header
// "fr1;fr2;fr3 --> R0;<R0 values>", one line per frequency triplet with its
// N_R0 thresholds, and a closing "time (s);<seconds>" line, numbers in the
// default stream format (6 significant digits, as printf %g).
extern "C" int bd_gilmore_sweep_write(const char* path, const double* fr_combo, int n_combo,
                                      const double* results, int n_r0, double min_r0, double step_r0,
                                      double seconds) {
    if (!path || !fr_combo || !results || n_combo <= 0 || n_r0 <= 0) return BD_ERR_INVALID;
    FILE* f = std::fopen(path, "w");
    if (!f) return BD_ERR_IO;
    std::fprintf(f, "fr1;fr2;fr3 --> R0");
    for (int i = 0; i < n_r0; ++i) std::fprintf(f, ";%g", min_r0 + i * step_r0);
    for (int c = 0; c < n_combo; ++c) {
        std::fprintf(f, "\n%g;%g;%g", fr_combo[3 * c], fr_combo[3 * c + 1], fr_combo[3 * c + 2]);
        for (int i = 0; i < n_r0; ++i) std::fprintf(f, ";%g", results[(size_t)c * n_r0 + i]);
    }
    std::fprintf(f, "\ntime (s);%g", seconds);
    return std::fclose(f) == 0 ? BD_OK : BD_ERR_IO;
}
