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
#include "wv_march.h"
#include "wv_internal.h"
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <type_traits>

namespace bd {

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
    const double two_dt3 = 2.0 * std::pow(dt, 3);
    const double ct3 = std::pow(p->c_inf_tissue, 3);
    c.a1 = (T)((2.0 * dt * 2.0 * p->c1 * p->tau1) / (two_dt3 * (2.0 * dt + 3.0 * p->tau1) * ct3));
    c.b1 = (T)(p->tau1 / (2.0 * dt + 3.0 * p->tau1));
    c.a2 = (T)((2.0 * dt * 2.0 * p->c2 * p->tau2) / (two_dt3 * (2.0 * dt + 3.0 * p->tau2) * ct3));
    c.b2 = (T)(p->tau2 / (2.0 * dt + 3.0 * p->tau2));
    return c;
}

// ---------------------------------------------------------------------------
// source plane + boundary layers (one launch, no inter-thread dependency)
// ---------------------------------------------------------------------------
template <typename T>
struct WvFaceArgs {
    T* P; const T* Pm1; T* pmax;
    int Nxy, Nz, lo, iz_tissue, boundary;
    double kz_w, kz_t, kxy_w, kxy_t;          // Courant numbers c dt / h (W:100, W:111, W:117-119)
    int* tick; int tick_inc;                  // graph replay: advance the device step counter
};

// Z faces of the interior columns (W:101-113), innermost layer first; the
// source plane iz=2 must already hold p^{n+1}.  Used by the drop-in boundary
// entry; the plan fuses this work into the interior kernel.
template <typename T>
__global__ void __launch_bounds__(256) wv_zfaces(WvFaceArgs<T> a) {
    const int N = a.Nxy, lo = a.lo, nin = N - 3 - lo;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nin * nin) return;
    const int ix = lo + g % nin, iy = lo + g / nin;
    const size_t plane = (size_t)N * N, c = (size_t)iy * N + ix;
    T v1, v0, u[3];
    if (a.boundary == BD_BC_ZERO) {
        v1 = v0 = u[0] = u[1] = u[2] = T(0);
    } else {
        const T s2 = a.P[2 * plane + c];
        const T m0 = a.Pm1[c], m1 = a.Pm1[plane + c], m2 = a.Pm1[2 * plane + c];
        v1 = mur<T>(m2, s2, m1, a.kz_w);
        v0 = mur<T>(m1, v1, m0, a.kz_w);
        const size_t zt = (size_t)(a.Nz - 4) * plane + c;
        const T b = a.P[zt];
        const T n0 = a.Pm1[zt], n1 = a.Pm1[zt + plane], n2 = a.Pm1[zt + 2 * plane], n3 = a.Pm1[zt + 3 * plane];
        u[0] = mur<T>(n0, b, n1, a.kz_t);
        u[1] = mur<T>(n1, u[0], n2, a.kz_t);
        u[2] = mur<T>(n2, u[1], n3, a.kz_t);
    }
    auto put = [&](size_t i, T v) { a.P[i] = v; if (a.pmax && v > a.pmax[i]) a.pmax[i] = v; };
    put(plane + c, v1);
    put(c, v0);
    for (int k = 0; k < 3; ++k) put((size_t)(a.Nz - 3 + k) * plane + c, u[k]);
}

// X faces then Y faces (W:121-133), innermost layer first, every layer
// reading its inner neighbour's new value; the Z faces are already in P.
// Threads [0, nX) own one (iy, iz) row of the x layers (iy inside); threads
// [nX, nX+nY) own one (ix, iz) column of the y layers (all ix).  A y-layer
// thread in an x-layer column recomputes the x-layer value it starts from.
// Each thread reads everything it needs (new inside values, p^n, old pmax)
// before its first store, so the kernel costs one round of memory latency.
template <typename T>
__global__ void __launch_bounds__(256) wv_xyfaces(WvFaceArgs<T> a) {
    const int N = a.Nxy, Nz = a.Nz, lo = a.lo, nin = N - 3 - lo;
    const long long nX = (long long)nin * Nz, nY = (long long)N * Nz;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // the last face pass of a replayed graph advances the step counter that
    // the next replay's interior kernels read (no kernel of this pass reads it)
    if (a.tick && g == 0) *a.tick += a.tick_inc;
    if (g >= nX + nY) return;
    const size_t plane = (size_t)N * N;
    const bool zero = a.boundary == BD_BC_ZERO;
    const bool pm = a.pmax != nullptr;
    // three layers outward from an inside cell (old pc, new v), old values m1..m3
    auto chain = [&](T pc, T v, T m1, T m2, T m3, double k, T out[3]) {
        out[0] = zero ? T(0) : mur<T>(pc, v, m1, k);
        out[1] = zero ? T(0) : mur<T>(m1, out[0], m2, k);
        out[2] = zero ? T(0) : mur<T>(m2, out[1], m3, k);
    };
    if (g < nX) {
        const int iy = lo + (int)(g % nin), iz = (int)(g / nin);
        const double k = iz < a.iz_tissue ? a.kxy_w : a.kxy_t;
        const size_t r0 = (size_t)iz * plane + (size_t)iy * N;
        T bh, mh[4], qh[3], bl = T(0), ml[4] = {T(0), T(0), T(0), T(0)}, ql[3] = {T(0), T(0), T(0)};
        bh = a.P[r0 + N - 4];
#pragma unroll
        for (int j = 0; j < 4; ++j) mh[j] = a.Pm1[r0 + N - 4 + j];
#pragma unroll
        for (int j = 0; j < 3; ++j) qh[j] = pm ? a.pmax[r0 + N - 3 + j] : T(0);
        if (lo) {
            bl = a.P[r0 + 3];
#pragma unroll
            for (int j = 0; j < 4; ++j) ml[j] = a.Pm1[r0 + 3 - j];
#pragma unroll
            for (int j = 0; j < 3; ++j) ql[j] = pm ? a.pmax[r0 + 2 - j] : T(0);
        }
        T xh[3], xl[3];
        chain(mh[0], bh, mh[1], mh[2], mh[3], k, xh);
        chain(ml[0], bl, ml[1], ml[2], ml[3], k, xl);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            a.P[r0 + N - 3 + j] = xh[j];
            if (pm && xh[j] > qh[j]) a.pmax[r0 + N - 3 + j] = xh[j];
        }
        if (lo) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                a.P[r0 + 2 - j] = xl[j];
                if (pm && xl[j] > ql[j]) a.pmax[r0 + 2 - j] = xl[j];
            }
        }
        return;
    }
    const long long h = g - nX;
    const int ix = (int)(h % N), iz = (int)(h / N);
    const double k = iz < a.iz_tissue ? a.kxy_w : a.kxy_t;
    const size_t zo = (size_t)iz * plane;
    // value at (ix, iyb, iz) after the x faces, for iyb an inside row
    auto after_x = [&](int iyb) -> T {
        const size_t r0 = zo + (size_t)iyb * N;
        if (ix >= lo && ix <= N - 4) return a.P[r0 + ix];
        if (zero) return T(0);
        if (ix >= N - 3) {
            T v = a.P[r0 + N - 4];
            for (int x = N - 3; x <= ix; ++x) v = mur<T>(a.Pm1[r0 + x - 1], v, a.Pm1[r0 + x], k);
            return v;
        }
        T v = a.P[r0 + 3];
        for (int x = 2; x >= ix; --x) v = mur<T>(a.Pm1[r0 + x + 1], v, a.Pm1[r0 + x], k);
        return v;
    };
    const size_t c = zo + ix;
    T mh[4], qh[3], ml[4] = {T(0), T(0), T(0), T(0)}, ql[3] = {T(0), T(0), T(0)}, bl = T(0);
    const T bh = after_x(N - 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) mh[j] = a.Pm1[c + (size_t)(N - 4 + j) * N];
#pragma unroll
    for (int j = 0; j < 3; ++j) qh[j] = pm ? a.pmax[c + (size_t)(N - 3 + j) * N] : T(0);
    if (lo) {
        bl = after_x(3);
#pragma unroll
        for (int j = 0; j < 4; ++j) ml[j] = a.Pm1[c + (size_t)(3 - j) * N];
#pragma unroll
        for (int j = 0; j < 3; ++j) ql[j] = pm ? a.pmax[c + (size_t)(2 - j) * N] : T(0);
    }
    T yh[3], yl[3];
    chain(mh[0], bh, mh[1], mh[2], mh[3], k, yh);
    chain(ml[0], bl, ml[1], ml[2], ml[3], k, yl);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        a.P[c + (size_t)(N - 3 + j) * N] = yh[j];
        if (pm && yh[j] > qh[j]) a.pmax[c + (size_t)(N - 3 + j) * N] = yh[j];
    }
    if (lo) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            a.P[c + (size_t)(2 - j) * N] = yl[j];
            if (pm && yl[j] > ql[j]) a.pmax[c + (size_t)(2 - j) * N] = yl[j];
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
constexpr int kDefaultVariant = 0;   // 64 x 4 tile, one row per thread
constexpr int kDeepVariant = 2;      // 64 x 4 tile, two rows per thread, plane prefetch
constexpr int kDefaultBlocks = 1024;
// Tiles bd_westervelt_plan_prepare times on the bound state and chooses from
// (the first entry is the static choice of launch_shape's rule).  One MI355X,
// FP32 tone burst, us per step, variants 13 / 10 / 2 / 0 (64 x 4 tile; one
// row, three planes of loads in flight / one row, two / two rows, one / one
// row, none) with p^{n-3}, p^{n-4}, R^{n-2} non-temporal: 128^3 28.3 / 25.7 /
// 32.0 / 27.0, 160^3 53.0 / 50.2 / 49.7 / -, 192^3 69.7 / 66.7 / 68.4 / 71.3,
// 256^3 194.0 / 192.1 / 187.2 / -.  FP64 with the plan's cache policy,
// variants 0 / 10: 41.0 / 53.7 (128^3), 134.3 / 147.3 (192^3), 363.7 / 399.1
// (256^3).  No geometric rule separates the FP32 tiles, so the plan measures.
constexpr int kTuneF32[] = {13, 10, kDeepVariant, kDefaultVariant};
constexpr int kTuneF64[] = {kDefaultVariant, 10};
// The tile taken without a measurement (plans without prepare, the
// adjoint's forward recompute): one row per thread with three planes of
// loads in flight for FP32 (within 11% of the fastest above), one row without
// prefetch for FP64.
constexpr int kStaticF32 = 13, kStaticF64 = kDefaultVariant;

// The 12-field state does not fit in the last-level cache: the step kernel
// then loads the history levels with the non-temporal policy.
static bool streams_past_mall(int Nxy, int Nz, size_t elem_bytes) {
    return (size_t)12 * Nxy * Nxy * Nz * elem_bytes > kMallBytes;
}

// History-load cache policy without a measurement (0 cached, 1 every level
// non-temporal, 2 only p^{n-3}, p^{n-4}, R^{n-2}): non-temporal once the
// 12-field state outgrows the MALL; BD_WV_NT forces one.
static int nt_rule(int Nxy, int Nz, size_t elem_bytes) {
    if (const char* e = std::getenv("BD_WV_NT")) {
        const int x = std::atoi(e);
        if (x >= 0 && x <= 2) return x;
    }
    return streams_past_mall(Nxy, Nz, elem_bytes) ? 1 : 0;
}

static int check_params(const bd_westervelt_params* p, int dtype, int Nxy, int Nz) {
    if (!p || (dtype != BD_F32 && dtype != BD_F64)) return BD_ERR_INVALID;
    if (Nxy < 10 || Nz < 10) return BD_ERR_INVALID;
    // the step kernel addresses a plane and its z-neighbours with 32-bit byte
    // offsets from a per-plane buffer descriptor (up to 4 planes past it)
    if ((long long)Nxy * Nxy * 4 * (dtype == BD_F64 ? 8 : 4) > 0xFFFFFFFFLL) return BD_ERR_INVALID;
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

static WvGeom make_geom(const bd_westervelt_params* p, int Nxy, int Nz, int variant, int target_blocks) {
    WvGeom g;
    const WvVariant v = kWvVariants[variant];
    g.Nxy = Nxy; g.Nz = Nz;
    g.xlo = p->full_domain ? 3 : 0;
    g.iz_tissue = first_tissue_plane(p, Nz);
    const int nin = Nxy - 3 - g.xlo;
    g.ntx = (nin + v.tx - 1) / v.tx;
    g.nty = (nin + v.ty * v.ry - 1) / (v.ty * v.ry);
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
    g.obj_on = 0; g.ocx = g.ocy = g.ocz = 0; g.or2 = 0.0;
    const double dV = p->dxy * p->dxy * p->dz;
    g.escale[0] = dV / (p->rho * p->c_inf * p->c_inf);
    g.escale[1] = dV / (p->rho_tissue * p->c_inf_tissue * p->c_inf_tissue);
    return g;
}

// Variant / target block count for the interior kernel without a
// measurement (kStaticF32 / kStaticF64); BD_WV_VARIANT and BD_WV_BLOCKS
// override them (used by the tile-shape sweep).
static void launch_shape(const bd_westervelt_params* p, int Nxy, int Nz, bool f64, int* variant, int* blocks) {
    (void)p; (void)Nxy; (void)Nz;
    *variant = f64 ? kStaticF64 : kStaticF32;
    *blocks = kDefaultBlocks;
    if (const char* v = std::getenv("BD_WV_VARIANT")) {
        const int x = std::atoi(v);
        if (x >= 0 && x < kWvNumVariants) *variant = x;
    }
    if (const char* v = std::getenv("BD_WV_BLOCKS")) {
        const int x = std::atoi(v);
        if (x > 0) *blocks = x;
    }
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

template <typename T, int V>
static void launch_variant(const WvMarchArgs<T>& a, int step_arg, hipStream_t st) {
    constexpr WvVariant v = kWvVariants[V];
    if (a.nt == 1)
        hipLaunchKernelGGL((wv_march<T, v.tx, v.ty, v.ry, v.pf, 1>), dim3(a.g.nblocks),
                           dim3(v.tx * v.ty), 0, st, a, step_arg);
    else if (a.nt == 2)
        hipLaunchKernelGGL((wv_march<T, v.tx, v.ty, v.ry, v.pf, 2>), dim3(a.g.nblocks),
                           dim3(v.tx * v.ty), 0, st, a, step_arg);
    else
        hipLaunchKernelGGL((wv_march<T, v.tx, v.ty, v.ry, v.pf, 0>), dim3(a.g.nblocks),
                           dim3(v.tx * v.ty), 0, st, a, step_arg);
}

template <typename T>
static int launch_march(const WvMarchArgs<T>& a, int variant, int step_arg, hipStream_t st) {
    switch (variant) {
        case 0: launch_variant<T, 0>(a, step_arg, st); break;
        case 1: launch_variant<T, 1>(a, step_arg, st); break;
        case 2: launch_variant<T, 2>(a, step_arg, st); break;
        case 3: launch_variant<T, 3>(a, step_arg, st); break;
        case 4: launch_variant<T, 4>(a, step_arg, st); break;
        case 5: launch_variant<T, 5>(a, step_arg, st); break;
        case 6: launch_variant<T, 6>(a, step_arg, st); break;
        case 7: launch_variant<T, 7>(a, step_arg, st); break;
        case 8: launch_variant<T, 8>(a, step_arg, st); break;
        case 9: launch_variant<T, 9>(a, step_arg, st); break;
        case 10: launch_variant<T, 10>(a, step_arg, st); break;
        case 11: launch_variant<T, 11>(a, step_arg, st); break;
        case 12: launch_variant<T, 12>(a, step_arg, st); break;
        case 13: launch_variant<T, 13>(a, step_arg, st); break;
        default: return BD_ERR_INVALID;
    }
    BD_LAUNCH_CHECK();
    return BD_OK;
}

template <typename T>
static WvFaceArgs<T> make_faces(const bd_westervelt_params* p, int Nxy, int Nz, T* P, const T* Pm1, T* pmax) {
    WvFaceArgs<T> f;
    f.P = P; f.Pm1 = Pm1; f.pmax = pmax;
    f.Nxy = Nxy; f.Nz = Nz; f.lo = p->full_domain ? 3 : 0;
    f.iz_tissue = first_tissue_plane(p, Nz);
    f.boundary = p->boundary;
    f.kz_w = p->c_inf * p->dt / p->dz;
    f.kz_t = p->c_inf_tissue * p->dt / p->dz;
    f.kxy_w = p->c_inf * p->dt / p->dxy;
    f.kxy_t = p->c_inf_tissue * p->dt / p->dxy;
    f.tick = nullptr;
    f.tick_inc = 0;
    return f;
}

template <typename T>
static int launch_xyfaces(const WvFaceArgs<T>& f, hipStream_t st) {
    const int nin = f.Nxy - 3 - f.lo;
    const long long n = (long long)nin * f.Nz + (long long)f.Nxy * f.Nz;
    hipLaunchKernelGGL((wv_xyfaces<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

template <typename T>
static int launch_zfaces(const WvFaceArgs<T>& f, hipStream_t st) {
    const int nin = f.Nxy - 3 - f.lo;
    hipLaunchKernelGGL((wv_zfaces<T>), dim3((nin * nin + 255) / 256), dim3(256), 0, st, f);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

// Fills the march arguments that fuse the source plane and the z, x and y
// faces and choose the history-load cache policy (a.g must be set).
template <typename T>
static void set_fused_z(WvMarchArgs<T>& a, const bd_westervelt_params* p, int on, const double* amp,
                        const double* phase) {
    a.fuse_z = on;
    a.boundary = p->boundary;
    a.kz_w = p->c_inf * p->dt / p->dz;
    a.kz_t = p->c_inf_tissue * p->dt / p->dz;
    a.fuse_xy = on;
    a.nt = nt_rule(a.g.Nxy, a.g.Nz, sizeof(T));   // a plan's measured policy replaces it
    a.kxy_w = p->c_inf * p->dt / p->dxy;
    a.kxy_t = p->c_inf_tissue * p->dt / p->dxy;
    a.d_time = nullptr;
    a.t_next = 0.0;
    a.src = make_src(p, amp, phase);
}

template <typename T>
static int forward_step_t(const bd_westervelt_params* prm, int Nxy, int Nz, const WvStepPtrs& q, int step,
                          double t_next, const double* amp, const double* phase, double* epart,
                          const WvObjective* obj, hipStream_t st) {
    WvMarchArgs<T> a;
    a.pn = (T*)q.pn; a.p = (const T*)q.p; a.pm1 = (const T*)q.pm1; a.pm2 = (const T*)q.pm2;
    a.pm3 = (const T*)q.pm3; a.pm4 = (const T*)q.pm4;
    a.r1 = (T*)q.r1; a.r1m1 = (const T*)q.r1m1; a.r1m2 = (const T*)q.r1m2;
    a.r2 = (T*)q.r2; a.r2m1 = (const T*)q.r2m1; a.r2m2 = (const T*)q.r2m2;
    a.pmax = nullptr; a.pmax_mask = 0; a.epart = epart; a.focus = nullptr; a.d_step = nullptr; a.d_step_next = nullptr; a.focus_idx = -1;
    int variant, blocks;
    launch_shape(prm, Nxy, Nz, sizeof(T) == 8, &variant, &blocks);
    a.g = make_geom(prm, Nxy, Nz, variant, blocks);
    if (obj) { a.g.obj_on = 1; a.g.ocx = obj->cx; a.g.ocy = obj->cy; a.g.ocz = obj->cz; a.g.or2 = obj->r2; }
    a.c = make_coef<T>(prm);
    set_fused_z<T>(a, prm, 1, amp, phase);
    a.t_next = t_next;
    return launch_march<T>(a, variant, step, st);
}

int wv_forward_step(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz, const WvStepPtrs& q, int step,
                    double t_next, const double* amp, const double* phase, double* epart, const WvObjective* obj,
                    hipStream_t st) {
    if (dtype == BD_F32) return forward_step_t<float>(prm, Nxy, Nz, q, step, t_next, amp, phase, epart, obj, st);
    return forward_step_t<double>(prm, Nxy, Nz, q, step, t_next, amp, phase, epart, obj, st);
}

int wv_forward_nblocks(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz) {
    int variant, blocks;
    launch_shape(prm, Nxy, Nz, dtype == BD_F64, &variant, &blocks);
    return make_geom(prm, Nxy, Nz, variant, blocks).nblocks;
}

int wv_check_params(const bd_westervelt_params* p, int dtype, int Nxy, int Nz) {
    return check_params(p, dtype, Nxy, Nz);
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
    a.pmax = nullptr; a.pmax_mask = 0; a.epart = nullptr; a.focus = nullptr; a.d_step = nullptr; a.d_step_next = nullptr; a.focus_idx = -1;
    int variant, blocks;
    launch_shape(prm, Nxy, Nz, sizeof(T) == 8, &variant, &blocks);
    a.g = make_geom(prm, Nxy, Nz, variant, blocks);
    a.c = make_coef<T>(prm);
    set_fused_z<T>(a, prm, 0, nullptr, nullptr);
    return launch_march<T>(a, variant, 0, st);
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
    (void)t;
    const WvFaceArgs<T> f = make_faces<T>(prm, Nxy, Nz, (T*)P, (const T*)Pm1, nullptr);
    int rc = launch_zfaces<T>(f, st);
    if (rc) return rc;
    return launch_xyfaces<T>(f, st);
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
    int variant = 0;
    void* state = nullptr; void* pmax = nullptr;
    double* energy = nullptr; double* focus = nullptr;
    const double* amp = nullptr; const double* phase = nullptr;
    int steps = 0;
    double* d_epart = nullptr;
    double* d_time = nullptr;
    int* d_step = nullptr;              // two step counters, one per graph of the pair
    std::vector<double> h_time;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    int device = 0;
    int target_blocks = 0;              // workgroup target of make_geom
    int epart_blocks = 0;               // per-step capacity of d_epart (largest candidate tile)
    bool tuned = false;                 // the step tile was chosen by measurement (prepare)
    int nt = 0;                         // history-load cache policy: nt_rule at creation, measured by prepare
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
    int blocks = 0;
    launch_shape(prm, Nxy, Nz, dtype == BD_F64, &pl->variant, &blocks);
    pl->nt = nt_rule(Nxy, Nz, dtype == BD_F64 ? 8 : 4);
    pl->target_blocks = blocks;
    pl->g = make_geom(prm, Nxy, Nz, pl->variant, blocks);
    pl->epart_blocks = pl->g.nblocks;
    for (int c : kTuneF32) pl->epart_blocks = std::max(pl->epart_blocks, make_geom(prm, Nxy, Nz, c, blocks).nblocks);
    for (int c : kTuneF64) pl->epart_blocks = std::max(pl->epart_blocks, make_geom(prm, Nxy, Nz, c, blocks).nblocks);
    const int lo = prm->full_domain ? 3 : 0;
    pl->focus_idx = -1;
    if (fx >= 0) {
        if (fx < lo || fx > Nxy - 4 || fy < lo || fy > Nxy - 4 || fz < 3 || fz > Nz - 4) { delete pl; return BD_ERR_INVALID; }
        pl->focus_idx = (long long)fz * Nxy * Nxy + (long long)fy * Nxy + fx;
    }
    if (hipGetDevice(&pl->device) != hipSuccess ||
        hipMalloc(&pl->d_epart, sizeof(double) * 2 * (size_t)pl->epart_blocks * max_steps) != hipSuccess ||
        hipMalloc(&pl->d_time, sizeof(double) * ((size_t)max_steps + 1)) != hipSuccess ||
        hipMalloc(&pl->d_step, 2 * sizeof(int)) != hipSuccess) {
        bd_westervelt_plan_destroy(pl);
        return BD_ERR_ALLOC;
    }
    // time table by repeated addition, as the reference's t = t + dt (W:564)
    std::vector<double>& tt = pl->h_time;
    tt.resize((size_t)max_steps + 1);
    double t = 0.0;
    for (int k = 0; k <= max_steps; ++k) { tt[k] = t; t = t + prm->dt; }
    if (hipMemcpy(pl->d_time, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(pl->d_epart, 0, sizeof(double) * 2 * (size_t)pl->epart_blocks * max_steps) != hipSuccess) {
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
    for (int k = 0; k < 2; ++k) {
        if (pl->gexec[k]) { (void)hipGraphExecDestroy(pl->gexec[k]); pl->gexec[k] = nullptr; }
        if (pl->graph[k]) { (void)hipGraphDestroy(pl->graph[k]); pl->graph[k] = nullptr; }
    }
    return BD_OK;
}

template <typename T>
static int plan_step(bd_westervelt_plan* pl, int n, const int* d_step, int step_arg, hipStream_t st,
                     int pmax_mask, int* d_step_next = nullptr) {
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
    a.pmax_mask = pmax_mask;
    a.epart = pl->energy ? pl->d_epart : nullptr;
    a.focus = pl->focus;
    a.d_step = d_step;
    a.d_step_next = d_step_next;
    a.focus_idx = pl->focus_idx;
    a.g = pl->g;
    a.c = make_coef<T>(&pl->prm);
    set_fused_z<T>(a, &pl->prm, 1, pl->amp, pl->phase);
    a.nt = pl->nt;
    if (d_step) a.d_time = pl->d_time;            // graph node: t = d_time[*d_step + j + 1]
    else a.t_next = pl->h_time[(size_t)n + 1];
    return launch_march<T>(a, pl->variant, step_arg, st);
}

// Interior pmax levels taken by the kernel of step n in a run that started at
// step s0: step n holds p^{n-4} .. p^{n+1}; the kernel whose new level is a
// multiple of 6 takes the levels of its window produced in this run.
static int pmax_window(int n, int s0) {
    if ((n + 1) % 6 != 0) return 0;
    return (1 << std::min(6, n + 1 - s0)) - 1;
}

// The same for a run ending at step `end`: its last step also takes the
// levels after the run's last window (at most five, all held by that step),
// so pmax is exact when the run's work is done, with no extra pass.
static int pmax_mask(int n, int s0, int end) {
    int m = pmax_window(n, s0);
    if (n == end - 1) {
        const int last = 6 * (end / 6);
        const int l0 = (last >= s0 + 1 ? last : s0) + 1;
        if (l0 <= end) m |= (1 << (end - l0 + 1)) - 1;
    }
    return m;
}

template <typename T>
static int plan_capture_t(bd_westervelt_plan* pl) {
    for (int k = 0; k < 2; ++k) {
        if (pl->gexec[k]) continue;
        hipStream_t cap;
        BD_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        BD_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        int rc = BD_OK;
        // replays start at a step n = 0 mod 6 inside the run, so the last
        // kernel's window (levels n+1 .. n+6) lies inside the run
        for (int j = 0; j < 6 && rc == BD_OK; ++j)
            rc = plan_step<T>(pl, j, pl->d_step + k, j, cap, j == 5 ? 63 : 0,
                              j == 5 ? pl->d_step + (1 - k) : nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return BD_ERR_HIP;
        pl->graph[k] = g;
        BD_HIP_CHECK(hipGraphInstantiate(&pl->gexec[k], g, nullptr, nullptr, 0));
    }
    return BD_OK;
}

template <typename T>
static int plan_run_t(bd_westervelt_plan* pl, int n_steps, int use_graph, hipStream_t st) {
    int n = pl->steps;
    const int s0 = n;
    const int end = n + n_steps;
    // eager steps until the ring phase is 0 mod 6
    while (n < end && (!use_graph || n % 6 != 0)) {
        int rc = plan_step<T>(pl, n, nullptr, n, st, pmax_mask(n, s0, end));
        if (rc) return rc;
        ++n;
        if (!use_graph) continue;
    }
    if (use_graph && end - n >= 6) {
        int rc = plan_capture_t<T>(pl);
        if (rc) return rc;
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, pl->d_step, n);
        BD_LAUNCH_CHECK();
        for (int k = 0; end - n >= 6; k ^= 1) {
            BD_HIP_CHECK(hipGraphLaunch(pl->gexec[k], st));
            n += 6;
        }
    }
    while (n < end) {
        int rc = plan_step<T>(pl, n, nullptr, n, st, pmax_mask(n, s0, end));
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


// Launches the step kernel of the next step again with tile `variant` and
// history cache policy `nt` (-1: the plan's) on geometry g (state not
// advanced; it rewrites the next levels, the step's energy partials and focus
// sample, which the real step overwrites, and with use_pmax the peak-pressure
// field).
template <typename T>
static int plan_rerun_step_t(bd_westervelt_plan* pl, hipStream_t st, int variant, const WvGeom& g, bool use_pmax,
                             int nt = -1) {
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
    a.pmax = use_pmax ? (T*)pl->pmax : nullptr;
    a.pmax_mask = (n + 1) % 6 == 0 ? 63 : 0;      // the window of the step inside a replay
    a.epart = pl->energy ? pl->d_epart : nullptr;
    a.focus = pl->focus;
    a.d_step = nullptr;
    a.d_step_next = nullptr;
    a.focus_idx = pl->focus_idx;
    a.g = g;
    a.c = make_coef<T>(&pl->prm);
    set_fused_z<T>(a, &pl->prm, 1, pl->amp, pl->phase);
    if (nt >= 0) a.nt = nt;
    else a.nt = pl->nt;
    a.t_next = pl->h_time[(size_t)n + 1];
    return launch_march<T>(a, variant, n, st);
}

extern "C" int bd_westervelt_plan_rerun_step(bd_westervelt_plan* pl, void* stream) {
    if (!pl || !pl->state || pl->steps >= pl->max_steps) return BD_ERR_INVALID;
    if (pl->dtype == BD_F32) return plan_rerun_step_t<float>(pl, as_stream(stream), pl->variant, pl->g, true);
    return plan_rerun_step_t<double>(pl, as_stream(stream), pl->variant, pl->g, true);
}

// Chooses the step tile and the history cache policy by measurement: each
// candidate relaunches the next step (2 untimed + 12 timed launches, pmax
// untouched, nothing the real step does not overwrite) on a private stream;
// the fastest mean is kept.  Policies: with a state past the last-level
// cache every history level non-temporal or only p^{n-3}, p^{n-4} and
// R^{n-2} (FP32 192^3, 340 MB: 66.9 / 66.0 us per step; 256^3, 805 MB:
// 163 / 186); with a state it holds, cached or the same partial policy.
// Runs once per plan, before its first step; a BD_WV_VARIANT or BD_WV_NT
// override skips it.
template <typename T>
static int plan_tune_t(bd_westervelt_plan* pl) {
    pl->tuned = true;
    if (std::getenv("BD_WV_VARIANT") || std::getenv("BD_WV_NT") || pl->steps != 0 || pl->steps >= pl->max_steps)
        return BD_OK;
    const bool f32 = std::is_same<T, float>::value;
    const int* cands = f32 ? kTuneF32 : kTuneF64;
    const int nc = f32 ? (int)(sizeof(kTuneF32) / sizeof(int)) : (int)(sizeof(kTuneF64) / sizeof(int));
    struct Res {
        hipStream_t st = nullptr; hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Res() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        }
    } r;
    BD_HIP_CHECK(hipStreamCreateWithFlags(&r.st, hipStreamNonBlocking));
    BD_HIP_CHECK(hipEventCreate(&r.e0));
    BD_HIP_CHECK(hipEventCreate(&r.e1));
    constexpr int kWarm = 2, kTimed = 12;
    const int pol[2] = {streams_past_mall(pl->Nxy, pl->Nz, sizeof(T)) ? 1 : 0, 2};
    int best = pl->variant, best_nt = pol[0];
    float best_ms = 1e30f;
    for (int i = 0; i < nc; ++i) {
        const WvGeom g = make_geom(&pl->prm, pl->Nxy, pl->Nz, cands[i], pl->target_blocks);
        if (g.nblocks > pl->epart_blocks) continue;
        for (const int nt : pol) {
            for (int k = 0; k < kWarm + kTimed; ++k) {
                if (k == kWarm) BD_HIP_CHECK(hipEventRecord(r.e0, r.st));
                const int rc = plan_rerun_step_t<T>(pl, r.st, cands[i], g, false, nt);
                if (rc) return rc;
            }
            BD_HIP_CHECK(hipEventRecord(r.e1, r.st));
            BD_HIP_CHECK(hipEventSynchronize(r.e1));
            float ms = 0.f;
            BD_HIP_CHECK(hipEventElapsedTime(&ms, r.e0, r.e1));
            if (ms < best_ms) { best_ms = ms; best = cands[i]; best_nt = nt; }
        }
    }
    pl->variant = best;
    pl->nt = best_nt;
    pl->g = make_geom(&pl->prm, pl->Nxy, pl->Nz, best, pl->target_blocks);
    return BD_OK;
}

extern "C" int bd_westervelt_plan_prepare(bd_westervelt_plan* pl) {
    if (!pl || !pl->state) return BD_ERR_INVALID;
    int dev = 0;
    BD_HIP_CHECK(hipGetDevice(&dev));
    if (dev != pl->device) return BD_ERR_INVALID;
    if (pl->dtype == BD_F32) {
        if (!pl->tuned && !pl->gexec[0]) { const int rc = plan_tune_t<float>(pl); if (rc) return rc; }
        return plan_capture_t<float>(pl);
    }
    if (!pl->tuned && !pl->gexec[0]) { const int rc = plan_tune_t<double>(pl); if (rc) return rc; }
    return plan_capture_t<double>(pl);
}

extern "C" long long bd_westervelt_plan_interior_cells(const bd_westervelt_plan* pl) {
    if (!pl) return -1;
    const long long nin = pl->Nxy - 3 - (pl->prm.full_domain ? 3 : 0);
    return nin * nin * (long long)(pl->Nz - 6);
}

extern "C" int bd_westervelt_plan_shape(const bd_westervelt_plan* pl, int* tile_x, int* tile_y, int* rows,
                                        int* prefetch, int* nontemporal, int* workgroups) {
    if (!pl) return BD_ERR_INVALID;
    const WvVariant v = kWvVariants[pl->variant];
    if (tile_x) *tile_x = v.tx;
    if (tile_y) *tile_y = v.ty;
    if (rows) *rows = v.ry;
    if (prefetch) *prefetch = v.pf;
    if (nontemporal) *nontemporal = pl->nt;
    if (workgroups) *workgroups = pl->g.nblocks;
    return BD_OK;
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

extern "C" int bd_westervelt_first_tissue_plane(const bd_westervelt_params* prm, int Nz) {
    if (!prm || Nz <= 0) return -1;
    return first_tissue_plane(prm, Nz);
}

extern "C" int bd_westervelt_plan_seek(bd_westervelt_plan* pl, int step) {
    if (!pl || !pl->state || step < 0 || step > pl->max_steps) return BD_ERR_INVALID;
    pl->steps = step;
    return BD_OK;
}

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
    for (int k = 0; k < 2; ++k) {
        if (pl->gexec[k]) (void)hipGraphExecDestroy(pl->gexec[k]);
        if (pl->graph[k]) (void)hipGraphDestroy(pl->graph[k]);
    }
    if (pl->d_epart) (void)hipFree(pl->d_epart);
    if (pl->d_time) (void)hipFree(pl->d_time);
    if (pl->d_step) (void)hipFree(pl->d_step);
    delete pl;
    return BD_OK;
}
