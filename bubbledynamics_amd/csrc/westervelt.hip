// Westervelt acoustics on gfx950: interior z-march, boundary/source kernel,
// water heat source, harmonic accumulators, and the run-level plan.
//
// Reference: Westervelt_relax_fft_Pressure_Energy.cu (W:line).
//
// Interior kernel (wv_march): one workgroup owns a TX x TY tile of columns and
// a chunk of z planes.  The centre plane p^n(z) is staged in LDS with a 3-cell
// x/y halo (double-buffered, one barrier per plane); the z-neighbours come from
// a 7-deep register queue fed by one coalesced load per plane.  The history
// levels p^{n-1..n-4}, the relaxation levels and pmax are streamed once.  Peak
// pressure, per-step energy and the focus sample are fused reductions.
#include "common.h"
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <type_traits>

namespace bd {

// ---------------------------------------------------------------------------
// coefficients (host-derived from the parameter struct, W:214-230)
// ---------------------------------------------------------------------------
template <typename T>
struct WvCoef {
    T kxy[2], kz[2], dlt[2], bet[2];   // [0] water, [1] tissue
    T cdt2;                            // (c_t dt)^2                       W:230
    T K1, den1, den1b, tau1;           // R1 recurrence                   W:227
    T K2, den2, den2b, tau2;           // R2 recurrence                   W:228
    T two_dt3;                         // 2 dt^3                          W:226
};

struct WvGeom {
    int Nxy, Nz, xlo, iz_tissue;
    int ntx, nty, nzc, zchunk, nblocks;
    int sponge_w, full_domain;
    double sponge_alpha;
    double escale[2];                  // dV / (rho c^2), water / tissue
};

template <typename T>
struct WvMarchArgs {
    T* pn; const T* p; const T* pm1; const T* pm2; const T* pm3; const T* pm4;
    T* r1; const T* r1m1; const T* r1m2; T* r2; const T* r2m1; const T* r2m2;
    T* pmax;            // nullable
    double* epart;      // [step][nblocks][2], nullable
    double* focus;      // [step], nullable
    const int* d_step;  // step base for graph replay, nullable
    long long focus_idx;
    WvGeom g;
    WvCoef<T> c;
};

static int first_tissue_plane(const bd_westervelt_params* p, int Nz) {
    for (int iz = 0; iz < Nz; ++iz)
        if (!(p->z_min + iz * p->dz < p->water_tissue)) return iz;
    return Nz;
}

template <typename T>
static WvCoef<T> make_coef(const bd_westervelt_params* p) {
    WvCoef<T> c;
    const double dt = p->dt;
    c.kxy[0] = (T)std::pow(p->c_inf * dt / p->dxy, 2);
    c.kz[0] = (T)std::pow(p->c_inf * dt / p->dz, 2);
    c.dlt[0] = (T)(p->delta / (2.0 * dt * std::pow(p->c_inf, 2)));
    c.bet[0] = (T)(2.0 * p->beta / (p->rho * std::pow(p->c_inf, 2)));
    c.kxy[1] = (T)std::pow(p->c_inf_tissue * dt / p->dxy, 2);
    c.kz[1] = (T)std::pow(p->c_inf_tissue * dt / p->dz, 2);
    c.dlt[1] = (T)(p->delta_tissue / (2.0 * dt * std::pow(p->c_inf_tissue, 2)));
    c.bet[1] = (T)(2.0 * p->beta_tissue / (p->rho_tissue * std::pow(p->c_inf_tissue, 2)));
    c.cdt2 = (T)std::pow(p->c_inf_tissue * dt, 2);
    c.K1 = (T)(2.0 * dt * 2.0 * p->c1 * p->tau1);
    c.den1 = (T)((2.0 * dt + 3.0 * p->tau1) * std::pow(p->c_inf_tissue, 3));
    c.den1b = (T)(2.0 * dt + 3.0 * p->tau1);
    c.tau1 = (T)p->tau1;
    c.K2 = (T)(2.0 * dt * 2.0 * p->c2 * p->tau2);
    c.den2 = (T)((2.0 * dt + 3.0 * p->tau2) * std::pow(p->c_inf_tissue, 3));
    c.den2b = (T)(2.0 * dt + 3.0 * p->tau2);
    c.tau2 = (T)p->tau2;
    c.two_dt3 = (T)(2.0 * std::pow(dt, 3));
    return c;
}

// ---------------------------------------------------------------------------
// interior z-march
// ---------------------------------------------------------------------------
template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
wv_march(WvMarchArgs<T> a, int step_arg) {
    constexpr int NT = TX * TY;
    __shared__ T s[2][TY + 6][TX + 6];
    __shared__ double red[NT / kWave];

    const int step = a.d_step ? (*a.d_step + step_arg) : step_arg;
    const int lb = xcd_remap((int)blockIdx.x, a.g.nblocks);
    const int zc = lb % a.g.nzc, t2 = lb / a.g.nzc;
    const int bx = t2 % a.g.ntx, by = t2 / a.g.ntx;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int N = a.g.Nxy;
    const size_t plane = (size_t)N * N;
    const int ix = a.g.xlo + bx * TX + tx, iy = a.g.xlo + by * TY + ty;
    const bool in_arr = ix < N && iy < N;
    const bool active = ix <= N - 4 && iy <= N - 4;
    const size_t cidx = (size_t)iy * N + ix;

    // halo sources of this thread (even mirror at index < 0, W:176-200)
    int gxl = a.g.xlo + bx * TX - 3 + tx; if (gxl < 0) gxl = -gxl;
    int gyb = a.g.xlo + by * TY - 3 + ty; if (gyb < 0) gyb = -gyb;
    const bool vL = tx < 3 && iy < N && gxl < N;
    const bool vR = tx >= TX - 3 && iy < N && ix + 3 < N;
    const bool vB = ty < 3 && ix < N && gyb < N;
    const bool vT = ty >= TY - 3 && ix < N && iy + 3 < N;
    const size_t offL = (size_t)iy * N + gxl, offR = cidx + 3;
    const size_t offB = (size_t)gyb * N + ix, offT = cidx + 3 * (size_t)N;

    const int z0 = 3 + zc * a.g.zchunk;
    const int z1 = min(z0 + a.g.zchunk, a.g.Nz - 3);

    T q[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = T(0);
    if (in_arr) {
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = a.p[(size_t)(z0 - 3 + k) * plane + cidx];
    }
    double e0 = 0.0, e1 = 0.0;
    const int W = a.g.sponge_w;

    for (int z = z0; z < z1; ++z) {
        const int b = (z - z0) & 1;
        const size_t zo = (size_t)z * plane;
        if (in_arr) q[6] = a.p[zo + 3 * plane + cidx];
        s[b][ty + 3][tx + 3] = q[3];
        if (tx < 3) s[b][ty + 3][tx] = vL ? a.p[zo + offL] : T(0);
        if (tx >= TX - 3) s[b][ty + 3][tx + 6] = vR ? a.p[zo + offR] : T(0);
        if (ty < 3) s[b][ty][tx + 3] = vB ? a.p[zo + offB] : T(0);
        if (ty >= TY - 3) s[b][ty + 6][tx + 3] = vT ? a.p[zo + offT] : T(0);

        const size_t idx = zo + cidx;
        const int reg = z >= a.g.iz_tissue ? 1 : 0;
        T pm1 = T(0), pm2 = T(0), pm3 = T(0), pm4 = T(0);
        T r1m1 = T(0), r1m2 = T(0), r2m1 = T(0), r2m2 = T(0), pmx = T(0);
        if (active) {
            pm1 = a.pm1[idx]; pm2 = a.pm2[idx]; pm3 = a.pm3[idx]; pm4 = a.pm4[idx];
            if (reg) { r1m1 = a.r1m1[idx]; r1m2 = a.r1m2[idx]; r2m1 = a.r2m1[idx]; r2m2 = a.r2m2[idx]; }
            if (a.pmax) pmx = a.pmax[idx];
        }
        __syncthreads();
        if (active) {
            const T pc = q[3];
            const T* row = &s[b][ty + 3][tx + 3];
            const T Dxx = (T(-490.0) * pc + T(270.0) * (row[1] + row[-1]) - T(27.0) * (row[2] + row[-2]) +
                           T(2.0) * (row[3] + row[-3])) / T(180.0);
            const T Dyy = (T(-490.0) * pc + T(270.0) * (s[b][ty + 4][tx + 3] + s[b][ty + 2][tx + 3]) -
                           T(27.0) * (s[b][ty + 5][tx + 3] + s[b][ty + 1][tx + 3]) +
                           T(2.0) * (s[b][ty + 6][tx + 3] + s[b][ty][tx + 3])) / T(180.0);
            const T Dzz = (T(-490.0) * pc + T(270.0) * (q[4] + q[2]) - T(27.0) * (q[5] + q[1]) +
                           T(2.0) * (q[6] + q[0])) / T(180.0);
            T Dttt, N12;
            if constexpr (std::is_same<T, float>::value) {
                // Same polynomials written on first differences of the history,
                // which FP32 forms exactly (Sterbenz); cuts cancellation ~10x.
                const T d1 = pc - pm1, d2 = pm1 - pm2, d3 = pm2 - pm3, d4 = pm3 - pm4;
                Dttt = T(5.0) * d1 - T(13.0) * d2 + T(11.0) * d3 - T(3.0) * d4;
                const T A = T(3.0) * d1 - d2;
                N12 = (A * A) / T(4.0) + pc * (T(2.0) * d1 - T(3.0) * d2 + d3);
            } else {
                // W:207-209, reference expression order
                Dttt = T(5.0) * pc - T(18.0) * pm1 + T(24.0) * pm2 - T(14.0) * pm3 + T(3.0) * pm4;
                const T A = T(3.0) * pc - T(4.0) * pm1 + pm2;
                N12 = (A * A) / T(4.0) + pc * (T(2.0) * pc - T(5.0) * pm1 + T(4.0) * pm2 - pm3);
            }
            T sumR = T(0);
            T r1 = T(0), r2 = T(0);
            if (reg) {  // W:226-230
                const T Dttt_2dt = Dttt / a.c.two_dt3;
                r1 = (a.c.K1 * Dttt_2dt / a.c.den1) + ((a.c.tau1 * (T(4.0) * r1m1 - r1m2)) / a.c.den1b);
                r2 = (a.c.K2 * Dttt_2dt / a.c.den2) + ((a.c.tau2 * (T(4.0) * r2m1 - r2m2)) / a.c.den2b);
                sumR = (r1 + r2) * a.c.cdt2;
            }
            const T lap = a.c.kxy[reg] * (Dxx + Dyy) + a.c.kz[reg] * Dzz;
            T pn;
            if constexpr (std::is_same<T, float>::value) {
                const T inc = (pc - pm1) + lap + a.c.dlt[reg] * Dttt + a.c.bet[reg] * N12 + sumR;
                pn = pc + inc;
            } else {  // W:234
                pn = T(2.0) * pc - pm1 + a.c.kxy[reg] * (Dxx + Dyy) + a.c.kz[reg] * Dzz +
                     a.c.dlt[reg] * Dttt + a.c.bet[reg] * N12 + sumR;
            }
            if (W > 0) {
                int dx = N - 1 - ix, dy = N - 1 - iy, dz = a.g.Nz - 1 - z;
                if (a.g.full_domain) { dx = min(dx, ix); dy = min(dy, iy); }
                T sp = T(0);
                if (dx < W) { const T u = T(a.g.sponge_alpha) * T(W - dx); sp += u * u; }
                if (dy < W) { const T u = T(a.g.sponge_alpha) * T(W - dy); sp += u * u; }
                if (dz < W) { const T u = T(a.g.sponge_alpha) * T(W - dz); sp += u * u; }
                if (sp != T(0)) pn = exp(-sp) * pn;
            }
            a.pn[idx] = pn;
            if (reg) { a.r1[idx] = r1; a.r2[idx] = r2; }
            if (a.pmax && pn > pmx) a.pmax[idx] = pn;
            if ((long long)idx == a.focus_idx && a.focus) a.focus[step] = (double)pn;
            const double dpn = (double)pn, dd = (double)pn - (double)pc;
            e0 += dpn * dpn * a.g.escale[reg];
            e1 += dd * dd - dpn * (double)lap;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = q[k + 1];
    }
    if (a.epart) {
        const double s0 = block_sum<NT>(e0, red);
        const double s1 = block_sum<NT>(e1, red);
        if (threadIdx.x == 0) {
            double* o = a.epart + ((size_t)step * a.g.nblocks + blockIdx.x) * 2;
            o[0] = s0;
            o[1] = s1;
        }
    }
}

// ---------------------------------------------------------------------------
// source plane + boundary layers (one launch, no inter-thread dependency)
// ---------------------------------------------------------------------------
struct WvSrc {
    int kind;                 // BD_SRC_*
    double xy_min, dxy, d, rmax, rmin, c_inf, fr, p0, burst_cycles, burst_t0;
    int arr_nx, arr_ny;
    double arr_x0, arr_y0, arr_pitch;
    const double* amp; const double* phase;
};

__device__ __forceinline__ double source_value(const WvSrc& s, int ix, int iy, double t) {
    const double X = s.xy_min + ix * s.dxy;
    const double Y = s.xy_min + iy * s.dxy;
    if (s.kind == BD_SRC_ARRAY) {
        const double fx = floor((X - s.arr_x0) / s.arr_pitch);
        const double fy = floor((Y - s.arr_y0) / s.arr_pitch);
        if (fx < 0 || fy < 0 || fx >= s.arr_nx || fy >= s.arr_ny) return 0.0;
        const int e = (int)fy * s.arr_nx + (int)fx;
        double tb = t;
        if (s.burst_cycles > 0) {
            tb = t - s.burst_t0;
            if (tb < 0.0 || tb * s.fr >= s.burst_cycles) return 0.0;
        }
        return s.amp[e] * sin(2 * M_PI * s.fr * tb + s.phase[e]);
    }
    // W:72-83
    const double r = sqrt(X * X + Y * Y);
    if (r <= s.rmax && r >= s.rmin) {
        const double rd = r / s.d;
        const double var = sqrt(1.0 + rd * rd);
        double tt = t + (s.d * (var - 1.0) / s.c_inf);
        if (s.burst_cycles > 0) {
            tt = tt - s.burst_t0;
            if (tt < 0.0 || tt * s.fr >= s.burst_cycles) return 0.0;
        }
        return s.p0 * sin(2 * M_PI * s.fr * tt) / var;
    }
    return 0.0;
}

template <typename T>
struct WvBndArgs {
    T* P; const T* Pm1; T* pmax;
    const int* d_step; const double* d_time;  // t = d_time[step+1] when d_step != null
    double t;                                 // otherwise
    int Nxy, Nz, lo, iz_tissue, boundary, fuse_source;
    double kz_w, kz_t, kxy_w, kxy_t;          // Courant numbers c dt / h (W:100, W:111, W:117-119)
    WvSrc src;
};

// Mur update in the expression order of W:104.
template <typename T>
__device__ __forceinline__ T mur(T pm1_nb, T p_nb, T pm1_self, double k) {
    return pm1_nb + ((p_nb - pm1_self) * T(k - 1.0) / T(k + 1.0));
}

template <typename T>
struct Bnd {
    const WvBndArgs<T>& a;
    double t;
    size_t plane;
    __device__ Bnd(const WvBndArgs<T>& a_, double t_) : a(a_), t(t_), plane((size_t)a_.Nxy * a_.Nxy) {}
    __device__ size_t at(int ix, int iy, int iz) const { return (size_t)iz * plane + (size_t)iy * a.Nxy + ix; }
    __device__ double kxy(int iz) const { return iz < a.iz_tissue ? a.kxy_w : a.kxy_t; }
    __device__ T src(int ix, int iy) const {
        return a.fuse_source ? (T)source_value(a.src, ix, iy, t) : a.P[at(ix, iy, 2)];
    }
    // New value of a cell in an interior column (x, y inside the update range).
    __device__ T val0(int ix, int iy, int iz) const {
        if (iz >= 3 && iz <= a.Nz - 4) return a.P[at(ix, iy, iz)];
        if (iz == 2) return src(ix, iy);
        if (a.boundary == BD_BC_ZERO) return T(0);
        if (iz <= 1) {
            T v = mur<T>(a.Pm1[at(ix, iy, 2)], src(ix, iy), a.Pm1[at(ix, iy, 1)], a.kz_w);
            if (iz == 1) return v;
            return mur<T>(a.Pm1[at(ix, iy, 1)], v, a.Pm1[at(ix, iy, 0)], a.kz_w);
        }
        T v = a.P[at(ix, iy, a.Nz - 4)];
        for (int z = a.Nz - 3; z <= iz; ++z) v = mur<T>(a.Pm1[at(ix, iy, z - 1)], v, a.Pm1[at(ix, iy, z)], a.kz_t);
        return v;
    }
    // New value at x-layer cell (ix in {0,1,2} or {N-3..N-1}), iy inside.
    __device__ T valX(int ix, int iy, int iz) const {
        if (a.boundary == BD_BC_ZERO) return T(0);
        const double k = kxy(iz);
        if (ix >= a.Nxy - 3) {
            T v = val0(a.Nxy - 4, iy, iz);
            for (int x = a.Nxy - 3; x <= ix; ++x) v = mur<T>(a.Pm1[at(x - 1, iy, iz)], v, a.Pm1[at(x, iy, iz)], k);
            return v;
        }
        T v = val0(3, iy, iz);
        for (int x = 2; x >= ix; --x) v = mur<T>(a.Pm1[at(x + 1, iy, iz)], v, a.Pm1[at(x, iy, iz)], k);
        return v;
    }
    __device__ T val1(int ix, int iy, int iz) const {
        const bool xl = ix >= a.Nxy - 3 || ix < a.lo;
        return xl ? valX(ix, iy, iz) : val0(ix, iy, iz);
    }
    __device__ void put(int ix, int iy, int iz, T v) const {
        const size_t i = at(ix, iy, iz);
        a.P[i] = v;
        if (a.pmax && v > a.pmax[i]) a.pmax[i] = v;
    }
};

template <typename T>
__global__ void __launch_bounds__(256) wv_boundary(WvBndArgs<T> a) {
    const double t = a.d_step ? a.d_time[*a.d_step + 1] : a.t;
    Bnd<T> B(a, t);
    const int N = a.Nxy, Nz = a.Nz, lo = a.lo;
    const int nin = N - 3 - lo;                    // interior columns per axis
    const long long nZ = (long long)nin * nin;
    const long long nX = (long long)nin * Nz;
    const long long nY = (long long)N * Nz;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nZ) {
        const int ix = lo + (int)(g % nin), iy = lo + (int)(g / nin);
        if (a.fuse_source) B.put(ix, iy, 2, B.src(ix, iy));
        B.put(ix, iy, 1, B.val0(ix, iy, 1));
        B.put(ix, iy, 0, B.val0(ix, iy, 0));
        for (int z = Nz - 3; z <= Nz - 1; ++z) B.put(ix, iy, z, B.val0(ix, iy, z));
        return;
    }
    if (g < nZ + nX) {
        const long long h = g - nZ;
        const int iy = lo + (int)(h % nin), iz = (int)(h / nin);
        // x-max layers, innermost first
        T v = B.val0(N - 4, iy, iz);
        const double k = B.kxy(iz);
        for (int x = N - 3; x <= N - 1; ++x) {
            v = (a.boundary == BD_BC_ZERO) ? T(0) : mur<T>(a.Pm1[B.at(x - 1, iy, iz)], v, a.Pm1[B.at(x, iy, iz)], k);
            B.put(x, iy, iz, v);
        }
        if (lo) {
            v = B.val0(3, iy, iz);
            for (int x = 2; x >= 0; --x) {
                v = (a.boundary == BD_BC_ZERO) ? T(0) : mur<T>(a.Pm1[B.at(x + 1, iy, iz)], v, a.Pm1[B.at(x, iy, iz)], k);
                B.put(x, iy, iz, v);
            }
        }
        return;
    }
    if (g < nZ + nX + nY) {
        const long long h = g - nZ - nX;
        const int ix = (int)(h % N), iz = (int)(h / N);
        const double k = B.kxy(iz);
        T v = B.val1(ix, N - 4, iz);
        for (int y = N - 3; y <= N - 1; ++y) {
            v = (a.boundary == BD_BC_ZERO) ? T(0) : mur<T>(a.Pm1[B.at(ix, y - 1, iz)], v, a.Pm1[B.at(ix, y, iz)], k);
            B.put(ix, y, iz, v);
        }
        if (lo) {
            v = B.val1(ix, 3, iz);
            for (int y = 2; y >= 0; --y) {
                v = (a.boundary == BD_BC_ZERO) ? T(0) : mur<T>(a.Pm1[B.at(ix, y + 1, iz)], v, a.Pm1[B.at(ix, y, iz)], k);
                B.put(ix, y, iz, v);
            }
        }
    }
}

// W:62-85: every column of plane iz=2.
template <typename T>
__global__ void __launch_bounds__(256) wv_source_kernel(T* P, int Nxy, double t, WvSrc src) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Nxy * Nxy) return;
    const int ix = g % Nxy, iy = g / Nxy;
    P[(size_t)2 * Nxy * Nxy + g] = (T)source_value(src, ix, iy, t);
}

__global__ void tick_kernel(int* d_step, int inc) { if (threadIdx.x == 0 && blockIdx.x == 0) *d_step += inc; }
__global__ void set_int_kernel(int* d, int v) { if (threadIdx.x == 0 && blockIdx.x == 0) *d = v; }

// W:242-265
template <typename T>
__global__ void water_power_kernel(T* Q, const T* Pnext, const T* Pm1, int Nxy, int nz_water,
                                   double dt, double scale, int last) {
    const size_t n = (size_t)Nxy * Nxy * nz_water;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!last) {
        const T dp = (Pnext[i] - Pm1[i]) / T(2.0 * dt);
        Q[i] = Q[i] + dp * dp * T(dt);
    } else {
        Q[i] = T(scale) * Q[i];
    }
}

// Running DFT bins (replaces W:268-321 + cuFFT).  The twiddles of the
// sample are formed once per workgroup in FP64 (exact integer phase m/N).
// With kahan=1 the buffer carries a compensation field per accumulator
// (fields 2*n_harm .. 4*n_harm-1) and the sums are compensated.
template <typename T>
__global__ void __launch_bounds__(256)
harmonics_kernel(T* H, const T* Pnext, size_t nf, size_t off, int i_time, int N_time, int periods,
                 int n_harm, int kahan) {
    __shared__ T tw[2][64];
    if (threadIdx.x < (unsigned)n_harm) {
        const int n = threadIdx.x + 1;
        const long long m = ((long long)n * periods * i_time) % N_time;
        const double ang = 2.0 * M_PI * (double)m / (double)N_time;
        tw[0][threadIdx.x] = (T)cos(ang);
        tw[1][threadIdx.x] = (T)sin(ang);
    }
    __syncthreads();
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nf) return;
    const T v = Pnext[off + j];
    for (int n = 0; n < n_harm; ++n) {
        T* Hr = H + (size_t)(2 * n) * nf;
        T* Hi = Hr + nf;
        const T xr = v * tw[0][n], xi = -(v * tw[1][n]);
        if (!kahan) {
            Hr[j] += xr;
            Hi[j] += xi;
        } else {
            T* Cr = H + (size_t)(2 * n_harm + 2 * n) * nf;
            T* Ci = Cr + nf;
            T y = xr - Cr[j], sum = Hr[j], t = sum + y;
            Cr[j] = (t - sum) - y; Hr[j] = t;
            y = xi - Ci[j]; sum = Hi[j]; t = sum + y;
            Ci[j] = (t - sum) - y; Hi[j] = t;
        }
    }
}

template <typename T>
__global__ void harmonics_q_kernel(T* Q, const T* H, size_t nf, size_t off, int N_time, int n_harm,
                                   double abs_coef, double fr, double rhoc, int kahan) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nf) return;
    double q = 0.0;
    for (int n = 1; n <= n_harm; ++n) {
        double re = (double)H[(size_t)(2 * (n - 1)) * nf + j];
        double im = (double)H[(size_t)(2 * (n - 1) + 1) * nf + j];
        if (kahan) {
            re -= (double)H[(size_t)(2 * n_harm + 2 * (n - 1)) * nf + j];
            im -= (double)H[(size_t)(2 * n_harm + 2 * (n - 1) + 1) * nf + j];
        }
        const double amp = sqrt(re * re + im * im) / N_time;
        const double In = 2.0 * (amp * amp) / rhoc;
        q = q + (2.0 * (abs_coef * n * fr / 1.0e6) * In);
    }
    Q[off + j] = (T)q;
}

__global__ void energy_finalize_kernel(const double* epart, double* energy, int step0, int nsteps,
                                       int nblocks) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsteps) return;
    const double* p = epart + (size_t)(step0 + s) * nblocks * 2;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nblocks; ++i) { a += p[2 * i]; b += p[2 * i + 1]; }
    energy[2 * (size_t)(step0 + s)] = a;
    energy[2 * (size_t)(step0 + s) + 1] = b;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kTX = 64, kTY = 4;

static int check_params(const bd_westervelt_params* p, int dtype, int Nxy, int Nz) {
    if (!p || (dtype != BD_F32 && dtype != BD_F64)) return BD_ERR_INVALID;
    if (Nxy < 10 || Nz < 10) return BD_ERR_INVALID;
    if (!(p->dt > 0) || !(p->dxy > 0) || !(p->dz > 0)) return BD_ERR_INVALID;
    if (p->boundary != BD_BC_MUR && p->boundary != BD_BC_ZERO) return BD_ERR_INVALID;
    if (p->source != BD_SRC_CAP && p->source != BD_SRC_ARRAY) return BD_ERR_INVALID;
    if (p->sponge_cells < 0 || p->sponge_cells > 256) return BD_ERR_INVALID;
    if (p->source == BD_SRC_CAP && !(p->d > 0)) return BD_ERR_INVALID;
    // 3-D leapfrog with the 6th-order Laplacian: stable iff c dt sqrt(sum 1/h^2) * sqrt(1088/180)/2 <= 1.
    const double cmax = std::max(p->c_inf, p->c_inf_tissue);
    const double s = cmax * p->dt * std::sqrt(2.0 / (p->dxy * p->dxy) + 1.0 / (p->dz * p->dz));
    if (s * std::sqrt(1088.0 / 180.0) / 2.0 > 1.0) return BD_ERR_UNSTABLE;
    return BD_OK;
}

static WvGeom make_geom(const bd_westervelt_params* p, int Nxy, int Nz, int target_blocks) {
    WvGeom g;
    g.Nxy = Nxy; g.Nz = Nz;
    g.xlo = p->full_domain ? 3 : 0;
    g.iz_tissue = first_tissue_plane(p, Nz);
    const int nin = Nxy - 3 - g.xlo;
    g.ntx = (nin + kTX - 1) / kTX;
    g.nty = (nin + kTY - 1) / kTY;
    const int nzin = Nz - 6;
    int nzc = std::max(1, target_blocks / (g.ntx * g.nty));
    int zchunk = (nzin + nzc - 1) / nzc;
    if (zchunk < 8) zchunk = std::min(8, nzin);
    nzc = (nzin + zchunk - 1) / zchunk;
    g.nzc = nzc; g.zchunk = zchunk;
    g.nblocks = g.ntx * g.nty * g.nzc;
    g.sponge_w = p->sponge_cells;
    g.full_domain = p->full_domain;
    g.sponge_alpha = p->sponge_alpha;
    const double dV = p->dxy * p->dxy * p->dz;
    g.escale[0] = dV / (p->rho * p->c_inf * p->c_inf);
    g.escale[1] = dV / (p->rho_tissue * p->c_inf_tissue * p->c_inf_tissue);
    return g;
}

static WvSrc make_src(const bd_westervelt_params* p, const double* amp, const double* phase) {
    WvSrc s;
    s.kind = p->source;
    s.xy_min = p->xy_min; s.dxy = p->dxy; s.d = p->d;
    s.rmax = p->d * std::tan(p->alpha);
    s.rmin = p->d * std::tan(p->alpha_hole);
    s.c_inf = p->c_inf; s.fr = p->fr; s.p0 = p->p0;
    s.burst_cycles = p->burst_cycles; s.burst_t0 = p->burst_t0;
    s.arr_nx = p->arr_nx; s.arr_ny = p->arr_ny;
    s.arr_x0 = p->arr_x0; s.arr_y0 = p->arr_y0; s.arr_pitch = p->arr_pitch;
    s.amp = amp; s.phase = phase;
    return s;
}

template <typename T>
static int launch_march(const WvMarchArgs<T>& a, int step_arg, hipStream_t st) {
    hipLaunchKernelGGL((wv_march<T, kTX, kTY>), dim3(a.g.nblocks), dim3(kTX * kTY), 0, st, a, step_arg);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

template <typename T>
static WvBndArgs<T> make_bnd(const bd_westervelt_params* p, int Nxy, int Nz, T* P, const T* Pm1,
                             T* pmax, const double* amp, const double* phase) {
    WvBndArgs<T> b;
    b.P = P; b.Pm1 = Pm1; b.pmax = pmax;
    b.d_step = nullptr; b.d_time = nullptr; b.t = 0.0;
    b.Nxy = Nxy; b.Nz = Nz; b.lo = p->full_domain ? 3 : 0;
    b.iz_tissue = first_tissue_plane(p, Nz);
    b.boundary = p->boundary;
    b.fuse_source = 1;
    b.kz_w = p->c_inf * p->dt / p->dz;
    b.kz_t = p->c_inf_tissue * p->dt / p->dz;
    b.kxy_w = p->c_inf * p->dt / p->dxy;
    b.kxy_t = p->c_inf_tissue * p->dt / p->dxy;
    b.src = make_src(p, amp, phase);
    return b;
}

template <typename T>
static int launch_bnd(const WvBndArgs<T>& b, hipStream_t st) {
    const int nin = b.Nxy - 3 - b.lo;
    const long long n = (long long)nin * nin + (long long)nin * b.Nz + (long long)b.Nxy * b.Nz;
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL((wv_boundary<T>), dim3(blocks), dim3(256), 0, st, b);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

}  // namespace bd

using namespace bd;

// ===========================================================================
// drop-in entry points
// ===========================================================================
template <typename T>
static int space6_t(const bd_westervelt_params* prm, void* Pnext, const void* P, const void* Pm1,
                    const void* Pm2, const void* Pm3, const void* Pm4, void* R1, const void* R1m1,
                    const void* R1m2, void* R2, const void* R2m1, const void* R2m2, int Nxy, int Nz,
                    hipStream_t st) {
    WvMarchArgs<T> a;
    a.pn = (T*)Pnext; a.p = (const T*)P; a.pm1 = (const T*)Pm1; a.pm2 = (const T*)Pm2;
    a.pm3 = (const T*)Pm3; a.pm4 = (const T*)Pm4;
    a.r1 = (T*)R1; a.r1m1 = (const T*)R1m1; a.r1m2 = (const T*)R1m2;
    a.r2 = (T*)R2; a.r2m1 = (const T*)R2m1; a.r2m2 = (const T*)R2m2;
    a.pmax = nullptr; a.epart = nullptr; a.focus = nullptr; a.d_step = nullptr; a.focus_idx = -1;
    a.g = make_geom(prm, Nxy, Nz, 2048);
    a.c = make_coef<T>(prm);
    return launch_march<T>(a, 0, st);
}

extern "C" int bd_westervelt_space6(const bd_westervelt_params* prm, int dtype, void* Pnext,
                                    const void* P, const void* Pm1, const void* Pm2, const void* Pm3,
                                    const void* Pm4, void* R1, const void* R1m1, const void* R1m2,
                                    void* R2, const void* R2m1, const void* R2m2, int Nxy, int Nz,
                                    void* stream) {
    int rc = check_params(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!Pnext || !P || !Pm1 || !Pm2 || !Pm3 || !Pm4 || !R1 || !R1m1 || !R1m2 || !R2 || !R2m1 || !R2m2)
        return BD_ERR_INVALID;
    if (dtype == BD_F32)
        return space6_t<float>(prm, Pnext, P, Pm1, Pm2, Pm3, Pm4, R1, R1m1, R1m2, R2, R2m1, R2m2, Nxy, Nz, as_stream(stream));
    return space6_t<double>(prm, Pnext, P, Pm1, Pm2, Pm3, Pm4, R1, R1m1, R1m2, R2, R2m1, R2m2, Nxy, Nz, as_stream(stream));
}

template <typename T>
static int source_t(const bd_westervelt_params* prm, void* P, double t, int Nxy, int Nz,
                    const double* amp, const double* phase, hipStream_t st) {
    (void)Nz;
    const WvSrc src = make_src(prm, amp, phase);
    hipLaunchKernelGGL((wv_source_kernel<T>), dim3((Nxy * Nxy + 255) / 256), dim3(256), 0, st, (T*)P, Nxy, t, src);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

extern "C" int bd_boundary_condition_source(const bd_westervelt_params* prm, int dtype, void* P,
                                            double t, int Nxy, int Nz, const double* elem_amp,
                                            const double* elem_phase, void* stream) {
    int rc = check_params(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!P) return BD_ERR_INVALID;
    if (prm->source == BD_SRC_ARRAY && (!elem_amp || !elem_phase || prm->arr_nx <= 0 || prm->arr_ny <= 0 || !(prm->arr_pitch > 0)))
        return BD_ERR_INVALID;
    if (dtype == BD_F32) return source_t<float>(prm, P, t, Nxy, Nz, elem_amp, elem_phase, as_stream(stream));
    return source_t<double>(prm, P, t, Nxy, Nz, elem_amp, elem_phase, as_stream(stream));
}

template <typename T>
static int abc_t(const bd_westervelt_params* prm, void* P, const void* Pm1, double t, int Nxy, int Nz,
                 hipStream_t st) {
    WvBndArgs<T> b = make_bnd<T>(prm, Nxy, Nz, (T*)P, (const T*)Pm1, nullptr, nullptr, nullptr);
    b.t = t; b.fuse_source = 0;
    return launch_bnd<T>(b, st);
}

extern "C" int bd_boundary_condition_space6(const bd_westervelt_params* prm, int dtype, void* P,
                                            const void* Pm1, double t, int Nxy, int Nz, void* stream) {
    int rc = check_params(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!P || !Pm1) return BD_ERR_INVALID;
    if (dtype == BD_F32) return abc_t<float>(prm, P, Pm1, t, Nxy, Nz, as_stream(stream));
    return abc_t<double>(prm, P, Pm1, t, Nxy, Nz, as_stream(stream));
}

extern "C" int bd_acoustic_power_water(const bd_westervelt_params* prm, int dtype, void* Q,
                                       const void* Pnext, const void* Pm1, int Nxy, int Nz, int last,
                                       void* stream) {
    int rc = check_params(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!Q || (!last && (!Pnext || !Pm1))) return BD_ERR_INVALID;
    const int nzw = first_tissue_plane(prm, Nz);
    const size_t n = (size_t)Nxy * Nxy * nzw;
    if (n == 0) return BD_OK;
    const double scale = 2.0 * prm->abs_water * prm->fr /
                         (std::pow(2.0 * M_PI * prm->fr, 2) * prm->rho * prm->c_inf);
    const int blocks = (int)((n + 255) / 256);
    hipStream_t st = as_stream(stream);
    if (dtype == BD_F32)
        hipLaunchKernelGGL((water_power_kernel<float>), dim3(blocks), dim3(256), 0, st, (float*)Q,
                           (const float*)Pnext, (const float*)Pm1, Nxy, nzw, prm->dt, scale, last);
    else
        hipLaunchKernelGGL((water_power_kernel<double>), dim3(blocks), dim3(256), 0, st, (double*)Q,
                           (const double*)Pnext, (const double*)Pm1, Nxy, nzw, prm->dt, scale, last);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

extern "C" int bd_harmonics_accumulate(int dtype, void* H, const void* Pnext, int Nxy, int Nz,
                                       int iz_tissue, int i_time, int N_time, int periods, int N_harm,
                                       int kahan, void* stream) {
    if ((dtype != BD_F32 && dtype != BD_F64) || !H || !Pnext) return BD_ERR_INVALID;
    if (Nxy <= 0 || Nz <= 0 || iz_tissue < 0 || iz_tissue >= Nz || N_time <= 0 || N_harm <= 0 ||
        N_harm > 64 || i_time < 0 || i_time >= N_time || periods <= 0)
        return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    const size_t plane = (size_t)Nxy * Nxy, nf = plane * (size_t)(Nz - iz_tissue);
    const int blocks = (int)((nf + 255) / 256);
    if (dtype == BD_F32)
        hipLaunchKernelGGL((harmonics_kernel<float>), dim3(blocks), dim3(256), 0, st, (float*)H,
                           (const float*)Pnext, nf, (size_t)iz_tissue * plane, i_time, N_time, periods,
                           N_harm, kahan ? 1 : 0);
    else
        hipLaunchKernelGGL((harmonics_kernel<double>), dim3(blocks), dim3(256), 0, st, (double*)H,
                           (const double*)Pnext, nf, (size_t)iz_tissue * plane, i_time, N_time, periods,
                           N_harm, kahan ? 1 : 0);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

extern "C" int bd_harmonics_q(const bd_westervelt_params* prm, int dtype, void* Q, const void* H,
                              int Nxy, int Nz, int iz_tissue, int N_time, int N_harm, double abs_coef,
                              int kahan, void* stream) {
    if (!prm || (dtype != BD_F32 && dtype != BD_F64) || !Q || !H) return BD_ERR_INVALID;
    if (Nxy <= 0 || Nz <= 0 || iz_tissue < 0 || iz_tissue >= Nz || N_time <= 0 || N_harm <= 0)
        return BD_ERR_INVALID;
    const size_t plane = (size_t)Nxy * Nxy, nf = plane * (size_t)(Nz - iz_tissue);
    const int blocks = (int)((nf + 255) / 256);
    const double rhoc = prm->rho_tissue * prm->c_inf_tissue;
    hipStream_t st = as_stream(stream);
    if (dtype == BD_F32)
        hipLaunchKernelGGL((harmonics_q_kernel<float>), dim3(blocks), dim3(256), 0, st, (float*)Q,
                           (const float*)H, nf, (size_t)iz_tissue * plane, N_time, N_harm, abs_coef,
                           prm->fr, rhoc, kahan ? 1 : 0);
    else
        hipLaunchKernelGGL((harmonics_q_kernel<double>), dim3(blocks), dim3(256), 0, st, (double*)Q,
                           (const double*)H, nf, (size_t)iz_tissue * plane, N_time, N_harm, abs_coef,
                           prm->fr, rhoc, kahan ? 1 : 0);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

// ===========================================================================
// run-level plan
// ===========================================================================
struct bd_westervelt_plan {
    bd_westervelt_params prm;
    int dtype, Nxy, Nz, max_steps;
    long long focus_idx;
    WvGeom g;
    void* state = nullptr; void* pmax = nullptr;
    double* energy = nullptr; double* focus = nullptr;
    const double* amp = nullptr; const double* phase = nullptr;
    int steps = 0;
    double* d_epart = nullptr;
    double* d_time = nullptr;
    int* d_step = nullptr;
    std::vector<double> h_time;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int device = 0;
};

extern "C" size_t bd_westervelt_state_elems(int Nxy, int Nz) {
    if (Nxy <= 0 || Nz <= 0) return 0;
    return (size_t)12 * Nxy * Nxy * Nz;
}

extern "C" int bd_westervelt_plan_create(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz,
                                         int max_steps, int fx, int fy, int fz,
                                         bd_westervelt_plan** out) {
    if (!out) return BD_ERR_INVALID;
    *out = nullptr;
    int rc = check_params(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (max_steps <= 0) return BD_ERR_INVALID;
    auto* pl = new bd_westervelt_plan();
    pl->prm = *prm; pl->dtype = dtype; pl->Nxy = Nxy; pl->Nz = Nz; pl->max_steps = max_steps;
    pl->g = make_geom(prm, Nxy, Nz, 2048);
    const int lo = prm->full_domain ? 3 : 0;
    pl->focus_idx = -1;
    if (fx >= 0) {
        if (fx < lo || fx > Nxy - 4 || fy < lo || fy > Nxy - 4 || fz < 3 || fz > Nz - 4) { delete pl; return BD_ERR_INVALID; }
        pl->focus_idx = (long long)fz * Nxy * Nxy + (long long)fy * Nxy + fx;
    }
    if (hipGetDevice(&pl->device) != hipSuccess ||
        hipMalloc(&pl->d_epart, sizeof(double) * 2 * (size_t)pl->g.nblocks * max_steps) != hipSuccess ||
        hipMalloc(&pl->d_time, sizeof(double) * ((size_t)max_steps + 1)) != hipSuccess ||
        hipMalloc(&pl->d_step, sizeof(int)) != hipSuccess) {
        bd_westervelt_plan_destroy(pl);
        return BD_ERR_ALLOC;
    }
    // time table by repeated addition, as the reference's t = t + dt (W:564)
    std::vector<double>& tt = pl->h_time;
    tt.resize((size_t)max_steps + 1);
    double t = 0.0;
    for (int k = 0; k <= max_steps; ++k) { tt[k] = t; t = t + prm->dt; }
    if (hipMemcpy(pl->d_time, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(pl->d_epart, 0, sizeof(double) * 2 * (size_t)pl->g.nblocks * max_steps) != hipSuccess) {
        bd_westervelt_plan_destroy(pl);
        return BD_ERR_HIP;
    }
    *out = pl;
    return BD_OK;
}

extern "C" int bd_westervelt_plan_bind(bd_westervelt_plan* pl, void* state, void* pmax, double* energy,
                                       double* focus, const double* amp, const double* phase) {
    if (!pl || !state) return BD_ERR_INVALID;
    if (pl->prm.source == BD_SRC_ARRAY && (!amp || !phase)) return BD_ERR_INVALID;
    pl->state = state; pl->pmax = pmax; pl->energy = energy; pl->focus = focus;
    pl->amp = amp; pl->phase = phase;
    pl->steps = 0;
    if (pl->gexec) { (void)hipGraphExecDestroy(pl->gexec); pl->gexec = nullptr; }
    if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
    return BD_OK;
}

template <typename T>
static int plan_step(bd_westervelt_plan* pl, int n, const int* d_step, int step_arg, hipStream_t st) {
    const size_t N = (size_t)pl->Nxy * pl->Nxy * pl->Nz;
    T* S = (T*)pl->state;
    auto P = [&](int m) { return S + (size_t)(((m % 6) + 6) % 6) * N; };
    auto R1 = [&](int m) { return S + (size_t)(6 + ((m % 3) + 3) % 3) * N; };
    auto R2 = [&](int m) { return S + (size_t)(9 + ((m % 3) + 3) % 3) * N; };
    WvMarchArgs<T> a;
    a.pn = P(n + 1); a.p = P(n); a.pm1 = P(n - 1); a.pm2 = P(n - 2); a.pm3 = P(n - 3); a.pm4 = P(n - 4);
    a.r1 = R1(n); a.r1m1 = R1(n - 1); a.r1m2 = R1(n - 2);
    a.r2 = R2(n); a.r2m1 = R2(n - 1); a.r2m2 = R2(n - 2);
    a.pmax = (T*)pl->pmax;
    a.epart = pl->energy ? pl->d_epart : nullptr;
    a.focus = pl->focus;
    a.d_step = d_step;
    a.focus_idx = pl->focus_idx;
    a.g = pl->g;
    a.c = make_coef<T>(&pl->prm);
    int rc = launch_march<T>(a, step_arg, st);
    if (rc) return rc;
    WvBndArgs<T> b = make_bnd<T>(&pl->prm, pl->Nxy, pl->Nz, P(n + 1), P(n), (T*)pl->pmax, pl->amp, pl->phase);
    b.d_step = d_step ? d_step : nullptr;
    b.d_time = pl->d_time;
    // eager steps pass the time directly; graph nodes read it from the table
    if (!d_step) {
        b.t = pl->h_time[(size_t)n + 1];
    } else {
        // the boundary kernel adds its node offset through d_time + offset
        b.d_time = pl->d_time + step_arg;
    }
    b.fuse_source = 1;
    return launch_bnd<T>(b, st);
}

template <typename T>
static int plan_run_t(bd_westervelt_plan* pl, int n_steps, int use_graph, hipStream_t st) {
    int n = pl->steps;
    const int end = n + n_steps;
    // eager steps until the ring phase is 0 mod 6
    while (n < end && (!use_graph || n % 6 != 0)) {
        int rc = plan_step<T>(pl, n, nullptr, n, st);
        if (rc) return rc;
        ++n;
        if (!use_graph) continue;
    }
    if (use_graph && end - n >= 6) {
        if (!pl->gexec) {
            hipStream_t cap;
            BD_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            BD_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            int rc = BD_OK;
            for (int j = 0; j < 6 && rc == BD_OK; ++j) rc = plan_step<T>(pl, j, pl->d_step, j, cap);
            if (rc == BD_OK) {
                hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(1), 0, cap, pl->d_step, 6);
            }
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(cap, &g);
            (void)hipStreamDestroy(cap);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return BD_ERR_HIP;
            pl->graph = g;
            BD_HIP_CHECK(hipGraphInstantiate(&pl->gexec, g, nullptr, nullptr, 0));
        }
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, pl->d_step, n);
        BD_LAUNCH_CHECK();
        while (end - n >= 6) {
            BD_HIP_CHECK(hipGraphLaunch(pl->gexec, st));
            n += 6;
        }
    }
    while (n < end) {
        int rc = plan_step<T>(pl, n, nullptr, n, st);
        if (rc) return rc;
        ++n;
    }
    pl->steps = n;
    return BD_OK;
}

extern "C" int bd_westervelt_plan_run(bd_westervelt_plan* pl, int n_steps, int use_graph, void* stream) {
    if (!pl || !pl->state || n_steps < 0) return BD_ERR_INVALID;
    if (pl->steps + n_steps > pl->max_steps) return BD_ERR_INVALID;
    if (n_steps == 0) return BD_OK;
    int dev = 0;
    BD_HIP_CHECK(hipGetDevice(&dev));
    if (dev != pl->device) return BD_ERR_INVALID;
    if (pl->dtype == BD_F32) return plan_run_t<float>(pl, n_steps, use_graph, as_stream(stream));
    return plan_run_t<double>(pl, n_steps, use_graph, as_stream(stream));
}

template <typename T>
static int plan_interior_only_t(bd_westervelt_plan* pl, hipStream_t st) {
    const int n = pl->steps;
    const size_t N = (size_t)pl->Nxy * pl->Nxy * pl->Nz;
    T* S = (T*)pl->state;
    auto P = [&](int m) { return S + (size_t)(((m % 6) + 6) % 6) * N; };
    auto R1 = [&](int m) { return S + (size_t)(6 + ((m % 3) + 3) % 3) * N; };
    auto R2 = [&](int m) { return S + (size_t)(9 + ((m % 3) + 3) % 3) * N; };
    WvMarchArgs<T> a;
    a.pn = P(n + 1); a.p = P(n); a.pm1 = P(n - 1); a.pm2 = P(n - 2); a.pm3 = P(n - 3); a.pm4 = P(n - 4);
    a.r1 = R1(n); a.r1m1 = R1(n - 1); a.r1m2 = R1(n - 2);
    a.r2 = R2(n); a.r2m1 = R2(n - 1); a.r2m2 = R2(n - 2);
    a.pmax = (T*)pl->pmax;
    a.epart = pl->energy ? pl->d_epart : nullptr;
    a.focus = pl->focus;
    a.d_step = nullptr;
    a.focus_idx = pl->focus_idx;
    a.g = pl->g;
    a.c = make_coef<T>(&pl->prm);
    return launch_march<T>(a, n, st);
}

extern "C" int bd_westervelt_plan_interior_only(bd_westervelt_plan* pl, void* stream) {
    if (!pl || !pl->state || pl->steps >= pl->max_steps) return BD_ERR_INVALID;
    if (pl->dtype == BD_F32) return plan_interior_only_t<float>(pl, as_stream(stream));
    return plan_interior_only_t<double>(pl, as_stream(stream));
}

extern "C" long long bd_westervelt_plan_interior_cells(const bd_westervelt_plan* pl) {
    if (!pl) return -1;
    const long long nin = pl->Nxy - 3 - (pl->prm.full_domain ? 3 : 0);
    return nin * nin * (long long)(pl->Nz - 6);
}

extern "C" int bd_westervelt_plan_finalize(bd_westervelt_plan* pl, void* stream) {
    if (!pl) return BD_ERR_INVALID;
    if (!pl->energy || pl->steps == 0) return BD_OK;
    hipStream_t st = as_stream(stream);
    const int blocks = (pl->steps + 127) / 128;
    hipLaunchKernelGGL(energy_finalize_kernel, dim3(blocks), dim3(128), 0, st, pl->d_epart, pl->energy, 0,
                       pl->steps, pl->g.nblocks);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

extern "C" int bd_westervelt_plan_steps(const bd_westervelt_plan* pl) { return pl ? pl->steps : -1; }

extern "C" int bd_westervelt_plan_slot(const bd_westervelt_plan* pl, int which, int level) {
    if (!pl || level < 0) return -1;
    const int n = pl->steps;
    if (which == 0) return (((n - level) % 6) + 6) % 6;
    if (which == 1) return 6 + (((n - 1 - level) % 3) + 3) % 3;
    if (which == 2) return 9 + (((n - 1 - level) % 3) + 3) % 3;
    return -1;
}

extern "C" int bd_westervelt_plan_destroy(bd_westervelt_plan* pl) {
    if (!pl) return BD_OK;
    if (pl->gexec) (void)hipGraphExecDestroy(pl->gexec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->d_epart) (void)hipFree(pl->d_epart);
    if (pl->d_time) (void)hipFree(pl->d_time);
    if (pl->d_step) (void)hipFree(pl->d_step);
    delete pl;
    return BD_OK;
}
