// Interior Westervelt update (W:139-236) as a 2.5-D z-march for gfx950.
//
// Workgroup = TX x TY threads; each thread owns RY columns (rows iy = base +
// r*TY + ty), so a workgroup covers a TX x (TY*RY) tile of columns and a chunk
// of z planes.  Per plane:
//   * p^n(z) of the tile plus a 3-cell x/y halo is staged in LDS (double
//     buffered: one barrier per plane);
//   * z-neighbours come from a 7-deep register queue per column fed by one
//     coalesced load of plane z+3;
//   * p^{n-1..n-4}, R1/R2 history and pmax are streamed once (when the state
//     is larger than the 256 MB last-level cache, every history level or only
//     the two oldest pressure levels and the older relaxation level with the
//     non-temporal cache policy, whichever the plan measures faster);
//   * with PF, all loads of plane z+1 are issued before plane z is computed.
// Fused outputs: pmax, per-block energy partials, the focus sample; with
// fuse_z the source plane and the z faces; with fuse_xy the x/y Mur faces,
// which a workgroup at the edge of the grid forms after its march as
// independent (plane, layer chain) tasks, every task's loads issued before
// any store so that the pass costs about one round of memory latency.
#pragma once
#include "common.h"
#include <type_traits>

namespace bd {

template <typename T>
struct WvCoef {
    T kxy[2], kz[2], dlt[2], bet[2];   // [0] water, [1] tissue
    T cdt2;                            // (c_t dt)^2                       W:230
    T K1, den1, den1b, tau1;           // R1 recurrence, reference form   W:227
    T K2, den2, den2b, tau2;           // R2 recurrence, reference form   W:228
    T two_dt3;                         // 2 dt^3                          W:226
    T a1, b1, a2, b2;                  // R = a*D3 + b*(4R' - R''), reciprocal form (FP32 path)
};

struct WvGeom {
    int Nxy, Nz, xlo, iz_tissue;
    int ntx, nty, nzc, zchunk, nblocks;
    int sponge_w, full_domain;
    double sponge_alpha;
    double escale[2];                  // dV / (rho c^2), water / tissue
    // objective mode: the two partials are sum p^2 inside / outside a sphere
    int obj_on, ocx, ocy, ocz;
    double or2;
};

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


// Mur update in the expression order of W:104.
template <typename T>
__device__ __forceinline__ T mur(T pm1_nb, T p_nb, T pm1_self, double k) {
    return pm1_nb + ((p_nb - pm1_self) * T(k - 1.0) / T(k + 1.0));
}

template <typename T>
struct WvMarchArgs {
    T* __restrict__ pn; const T* __restrict__ p;
    const T* __restrict__ pm1; const T* __restrict__ pm2; const T* __restrict__ pm3; const T* __restrict__ pm4;
    T* __restrict__ r1; const T* __restrict__ r1m1; const T* __restrict__ r1m2;
    T* __restrict__ r2; const T* __restrict__ r2m1; const T* __restrict__ r2m2;
    T* __restrict__ pmax;  // nullable
    // Interior pmax this step: 0 = not touched; else bit k takes level n+1-k
    // (k = 0..5: p^{n+1}, p^n, p^{n-1} .. p^{n-4}, all in registers here), so
    // one update every 6 steps covers every level.  Face cells update every step.
    int pmax_mask;
    double* epart;         // [step][nblocks][2], nullable
    double* focus;         // [step], nullable
    const int* d_step;     // step base for graph replay, nullable
    int* d_step_next;      // written with the next replay's base (last kernel of a graph), nullable
    long long focus_idx;
    WvGeom g;
    WvCoef<T> c;
    // source plane and z faces fused into the first / last z-chunk
    int fuse_z, boundary;
    double kz_w, kz_t;         // Courant numbers of the z faces (W:100, W:111)
    int fuse_xy;               // x/y faces after the march (requires fuse_z)
    int nt;                    // history loads: 0 cached, 1 all non-temporal, 2 p^{n-3}, p^{n-4}, R^{n-2} non-temporal
    double kxy_w, kxy_t;       // Courant numbers of the x/y faces (W:117-119)
    const double* d_time;      // time table (graph replay), t = d_time[step+1]
    double t_next;             // t^{n+1} when d_time is null
    WvSrc src;
};

// Compensated running sum (FP32 accumulators) / plain sum (FP64).
template <typename T>
struct Acc {
    T s = T(0), c = T(0);
    __device__ __forceinline__ void add(T v) {
        if constexpr (std::is_same<T, float>::value) {
            const T y = v - c, t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <typename T, int RY>
struct PlaneIn {
    T qn[RY];              // p^n at z+3
    T hl[RY], hr[RY];      // x halo (left / right), only for edge threads
    T hb, ht;              // y halo rows, only for edge threads
    T pm1[RY], pm2[RY], pm3[RY], pm4[RY];
    T r1m1[RY], r1m2[RY], r2m1[RY], r2m2[RY], pmx[RY];
};

// The FP64 one-row tile without prefetch is held to 128 VGPRs (4 waves per
// SIMD, 12 bytes per lane spilled outside the march; 130 VGPRs and 3 waves
// otherwise): 192^3 158 -> 135 us per step together with the loads above.
template <typename T, int TX, int TY, int RY, int PF, int NTM>
__global__ void __launch_bounds__(TX * TY, sizeof(T) == 8 && RY == 1 && PF == 0 ? 4 : 1)
wv_march(WvMarchArgs<T> a, int step_arg) {
    constexpr int NT = TX * TY;
    constexpr int ROWS = TY * RY;
    constexpr bool F32 = std::is_same<T, float>::value;
    // the three y-halo rows on each side are loaded one per thread row
    static_assert(TY >= 3 && TX >= 6, "tile too small for the 3-cell halo");
    __shared__ T s[2][ROWS + 6][TX + 6];
    __shared__ double red[NT / kWave];

    const int step = a.d_step ? (*a.d_step + step_arg) : step_arg;
    const bool pm_upd = a.pmax != nullptr && a.pmax_mask != 0;
    if (a.d_step_next && blockIdx.x == 0 && threadIdx.x == 0) *a.d_step_next = step + 1;
    const int lb = xcd_remap((int)blockIdx.x, a.g.nblocks);
    // tile-major: consecutive ids are neighbouring tiles of one z-chunk, whose
    // x/y halos overlap, so one XCD works on them together
    const int zc = lb / (a.g.ntx * a.g.nty), t2 = lb % (a.g.ntx * a.g.nty);
    const int bx = t2 % a.g.ntx, by = t2 / a.g.ntx;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int N = a.g.Nxy;
    const size_t plane = (size_t)N * N;
    const int ix = a.g.xlo + bx * TX + tx;
    const int ybase = a.g.xlo + by * ROWS;

    // In-plane offsets are 32-bit (a plane holds < 2^31 cells); plane base
    // pointers are workgroup-uniform, so loads use scalar-base addressing.
    int iyr[RY];
    bool inr[RY], act[RY];
    unsigned cidx[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        iyr[r] = ybase + r * TY + ty;
        inr[r] = ix < N && iyr[r] < N;
        act[r] = ix <= N - 4 && iyr[r] <= N - 4;
        cidx[r] = (unsigned)iyr[r] * (unsigned)N + (unsigned)ix;
    }
    // halo sources (even mirror at index < 0, W:176-200)
    int gxl = a.g.xlo + bx * TX - 3 + tx; if (gxl < 0) gxl = -gxl;
    const bool xL = tx < 3, xR = tx >= TX - 3;
    // x-halo offset relative to the thread's own column
    const int dxl = gxl - ix;
    const bool yB = ty < 3, yT = ty >= TY - 3;
    int gyb = ybase - 3 + ty; if (gyb < 0) gyb = -gyb;
    const int gyt = ybase + ROWS + (ty - (TY - 3));
    const unsigned offB = (unsigned)gyb * (unsigned)N + (unsigned)ix;
    const unsigned offT = (unsigned)gyt * (unsigned)N + (unsigned)ix;
    bool any_act = false;
#pragma unroll
    for (int r = 0; r < RY; ++r) any_act |= act[r];

    const int z0 = 3 + zc * a.g.zchunk;
    const int z1 = min(z0 + a.g.zchunk, a.g.Nz - 3);

    constexpr uint32_t ES = sizeof(T);
    const uint32_t pbytes = (uint32_t)(plane * ES);
    // Loads of plane z.  FP64 without prefetch: every load is issued whatever
    // the plane (a water plane's relaxation levels and the peak pressure
    // outside a window step read through zero-sized buffer descriptors: zeros,
    // no memory access), so the number of loads in flight does not depend on
    // the path taken and the compiler's waits stay exact.  FP32 measured
    // slower that way (71.1 -> 73.3 us per step at 192^3) and keeps the
    // branches, as do the prefetching tiles, whose plane buffers are copied at
    // the back-edge.
    auto load_plane = [&](int z, PlaneIn<T, RY>& v) {
        const size_t zo = (size_t)z * plane;
        const bool tis = z >= a.g.iz_tissue;
        // p from plane z to the end of the field (halo rows past the grid read
        // memory of the next plane or, past the field, zeros: never used)
        const rsrc_t rp = make_rsrc(a.p + zo, (size_t)(a.g.Nz - z) * plane * ES);
        if constexpr (!F32 && PF == 0) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const uint32_t o = cidx[r] * ES;
                v.hl[r] = xL ? bload<T>(rp, o + (uint32_t)(dxl * (int)ES)) : T(0);
                v.hr[r] = xR ? bload<T>(rp, o + 3 * ES) : T(0);
            }
            v.hb = yB ? bload<T>(rp, offB * ES) : T(0);
            v.ht = yT ? bload<T>(rp, offT * ES) : T(0);
            const rsrc_t q1 = make_rsrc(a.pm1 + zo, pbytes), q2 = make_rsrc(a.pm2 + zo, pbytes);
            const rsrc_t q3 = make_rsrc(a.pm3 + zo, pbytes), q4 = make_rsrc(a.pm4 + zo, pbytes);
            const uint32_t rb = tis ? pbytes : 0u;
            const rsrc_t s1 = make_rsrc(a.r1m1 + zo, rb), s2 = make_rsrc(a.r1m2 + zo, rb);
            const rsrc_t s3 = make_rsrc(a.r2m1 + zo, rb), s4 = make_rsrc(a.r2m2 + zo, rb);
            const rsrc_t sm = make_rsrc(pm_upd ? a.pmax + zo : a.p, pm_upd ? pbytes : 0u);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const uint32_t o = cidx[r] * ES;
                v.qn[r] = bload<T>(rp, o, 3 * pbytes);
                if constexpr (NTM == 1) {
                    v.pm1[r] = bload_nt<T>(q1, o); v.pm2[r] = bload_nt<T>(q2, o);
                    v.pm3[r] = bload_nt<T>(q3, o); v.pm4[r] = bload_nt<T>(q4, o);
                    v.r1m1[r] = bload_nt<T>(s1, o); v.r1m2[r] = bload_nt<T>(s2, o);
                    v.r2m1[r] = bload_nt<T>(s3, o); v.r2m2[r] = bload_nt<T>(s4, o);
                } else if constexpr (NTM == 2) {
                    v.pm1[r] = bload<T>(q1, o); v.pm2[r] = bload<T>(q2, o);
                    v.pm3[r] = bload_nt<T>(q3, o); v.pm4[r] = bload_nt<T>(q4, o);
                    v.r1m1[r] = bload<T>(s1, o); v.r1m2[r] = bload_nt<T>(s2, o);
                    v.r2m1[r] = bload<T>(s3, o); v.r2m2[r] = bload_nt<T>(s4, o);
                } else {
                    v.pm1[r] = bload<T>(q1, o); v.pm2[r] = bload<T>(q2, o);
                    v.pm3[r] = bload<T>(q3, o); v.pm4[r] = bload<T>(q4, o);
                    v.r1m1[r] = bload<T>(s1, o); v.r1m2[r] = bload<T>(s2, o);
                    v.r2m1[r] = bload<T>(s3, o); v.r2m2[r] = bload<T>(s4, o);
                }
                v.pmx[r] = bload<T>(sm, o);
            }
            return;
        }
        const rsrc_t q1 = make_rsrc(a.pm1 + zo, pbytes), q2 = make_rsrc(a.pm2 + zo, pbytes);
        const rsrc_t q3 = make_rsrc(a.pm3 + zo, pbytes), q4 = make_rsrc(a.pm4 + zo, pbytes);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const uint32_t o = cidx[r] * ES;
            v.qn[r] = bload<T>(rp, o, 3 * pbytes);
            v.hl[r] = xL ? bload<T>(rp, o + (uint32_t)(dxl * (int)ES)) : T(0);
            v.hr[r] = xR ? bload<T>(rp, o + 3 * ES) : T(0);
            if constexpr (NTM == 1) {
                v.pm1[r] = bload_nt<T>(q1, o); v.pm2[r] = bload_nt<T>(q2, o);
                v.pm3[r] = bload_nt<T>(q3, o); v.pm4[r] = bload_nt<T>(q4, o);
            } else if constexpr (NTM == 2) {
                v.pm1[r] = bload<T>(q1, o); v.pm2[r] = bload<T>(q2, o);
                v.pm3[r] = bload_nt<T>(q3, o); v.pm4[r] = bload_nt<T>(q4, o);
            } else {
                v.pm1[r] = bload<T>(q1, o); v.pm2[r] = bload<T>(q2, o);
                v.pm3[r] = bload<T>(q3, o); v.pm4[r] = bload<T>(q4, o);
            }
        }
        if (tis) {
            const rsrc_t s1 = make_rsrc(a.r1m1 + zo, pbytes), s2 = make_rsrc(a.r1m2 + zo, pbytes);
            const rsrc_t s3 = make_rsrc(a.r2m1 + zo, pbytes), s4 = make_rsrc(a.r2m2 + zo, pbytes);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const uint32_t o = cidx[r] * ES;
                if constexpr (NTM == 1) {
                    v.r1m1[r] = bload_nt<T>(s1, o); v.r1m2[r] = bload_nt<T>(s2, o);
                    v.r2m1[r] = bload_nt<T>(s3, o); v.r2m2[r] = bload_nt<T>(s4, o);
                } else if constexpr (NTM == 2) {
                    v.r1m1[r] = bload<T>(s1, o); v.r1m2[r] = bload_nt<T>(s2, o);
                    v.r2m1[r] = bload<T>(s3, o); v.r2m2[r] = bload_nt<T>(s4, o);
                } else {
                    v.r1m1[r] = bload<T>(s1, o); v.r1m2[r] = bload<T>(s2, o);
                    v.r2m1[r] = bload<T>(s3, o); v.r2m2[r] = bload<T>(s4, o);
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < RY; ++r) v.r1m1[r] = v.r1m2[r] = v.r2m1[r] = v.r2m2[r] = T(0);
        }
        if (pm_upd) {
            const rsrc_t sm = make_rsrc(a.pmax + zo, pbytes);
#pragma unroll
            for (int r = 0; r < RY; ++r) v.pmx[r] = bload<T>(sm, cidx[r] * ES);
        } else {
#pragma unroll
            for (int r = 0; r < RY; ++r) v.pmx[r] = T(0);
        }
        v.hb = yB ? bload<T>(rp, offB * ES) : T(0);
        v.ht = yT ? bload<T>(rp, offT * ES) : T(0);
    };

    T q[RY][7];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
        for (int k = 0; k < 7; ++k) q[r][k] = T(0);
        if (inr[r]) {
#pragma unroll
            for (int k = 0; k < 6; ++k) q[r][k] = (a.p + (size_t)(z0 - 3 + k) * plane)[cidx[r]];
        }
    }
    Acc<T> e0, e1;
    const int W = a.g.sponge_w;

    // Source plane (W:62-85) and z-min faces (W:101-105) of this tile's columns;
    // they depend only on the source and on p^n.
    if (a.fuse_z && zc == 0) {
        const double t = a.d_time ? a.d_time[step + 1] : a.t_next;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (!act[r]) continue;
            const T sv = (T)source_value(a.src, ix, iyr[r], t);
            T v1 = T(0), v0 = T(0);
            const size_t c = cidx[r];
            if (a.boundary != BD_BC_ZERO) {
                const T m0 = a.p[c], m1 = a.p[plane + c], m2 = a.p[2 * plane + c];
                v1 = mur<T>(m2, sv, m1, a.kz_w);
                v0 = mur<T>(m1, v1, m0, a.kz_w);
            }
            a.pn[2 * plane + c] = sv;
            a.pn[plane + c] = v1;
            a.pn[c] = v0;
            if (a.pmax) {
                if (sv > a.pmax[2 * plane + c]) a.pmax[2 * plane + c] = sv;
                if (v1 > a.pmax[plane + c]) a.pmax[plane + c] = v1;
                if (v0 > a.pmax[c]) a.pmax[c] = v0;
            }
        }
    }
    T lastp[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) lastp[r] = T(0);

    // PF planes of loads in flight ahead of the plane being computed
    static_assert(PF >= 0 && PF <= 3, "prefetch depth 0..3");
    PlaneIn<T, RY> cur, nx1, nx2;
    load_plane(z0, cur);
    if (PF >= 2 && z0 + 1 < z1) load_plane(z0 + 1, nx1);
    if (PF >= 3 && z0 + 2 < z1) load_plane(z0 + 2, nx2);
    for (int z = z0; z < z1; ++z) {
        PlaneIn<T, RY> nxt;
        if (PF >= 1 && z + PF < z1) load_plane(z + PF, nxt);
        const int b = (z - z0) & 1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            q[r][6] = cur.qn[r];
            const int ly = 3 + r * TY + ty;
            s[b][ly][tx + 3] = q[r][3];
            if (xL) s[b][ly][tx] = cur.hl[r];
            if (xR) s[b][ly][tx + 6] = cur.hr[r];
        }
        if (yB) s[b][ty][tx + 3] = cur.hb;
        if (yT) s[b][ROWS + 3 + (ty - (TY - 3))][tx + 3] = cur.ht;
        __syncthreads();
        const int reg = z >= a.g.iz_tissue ? 1 : 0;
        const size_t zo = (size_t)z * plane;
        if (any_act) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                if (!act[r]) continue;
                const int ly = 3 + r * TY + ty;
                const T pc = q[r][3];
                const T* row = &s[b][ly][tx + 3];
                const T sx1 = row[1] + row[-1], sx2 = row[2] + row[-2], sx3 = row[3] + row[-3];
                const T sy1 = s[b][ly + 1][tx + 3] + s[b][ly - 1][tx + 3];
                const T sy2 = s[b][ly + 2][tx + 3] + s[b][ly - 2][tx + 3];
                const T sy3 = s[b][ly + 3][tx + 3] + s[b][ly - 3][tx + 3];
                const T sz1 = q[r][4] + q[r][2], sz2 = q[r][5] + q[r][1], sz3 = q[r][6] + q[r][0];
                const T pm1 = cur.pm1[r], pm2 = cur.pm2[r], pm3 = cur.pm3[r], pm4 = cur.pm4[r];
                T Dxx, Dyy, Dzz, Dttt, N12, lap, pn, r1v = T(0), r2v = T(0);
                if constexpr (F32) {
                    constexpr T i180 = T(1.0 / 180.0);
                    Dxx = (T(-490.0) * pc + T(270.0) * sx1 - T(27.0) * sx2 + T(2.0) * sx3) * i180;
                    Dyy = (T(-490.0) * pc + T(270.0) * sy1 - T(27.0) * sy2 + T(2.0) * sy3) * i180;
                    Dzz = (T(-490.0) * pc + T(270.0) * sz1 - T(27.0) * sz2 + T(2.0) * sz3) * i180;
                    // the same polynomials on first differences of the history, which
                    // FP32 forms exactly (Sterbenz): ~10x less cancellation
                    const T d1 = pc - pm1, d2 = pm1 - pm2, d3 = pm2 - pm3, d4 = pm3 - pm4;
                    Dttt = T(5.0) * d1 - T(13.0) * d2 + T(11.0) * d3 - T(3.0) * d4;
                    const T A = T(3.0) * d1 - d2;
                    N12 = T(0.25) * (A * A) + pc * (T(2.0) * d1 - T(3.0) * d2 + d3);
                    T sumR = T(0);
                    if (reg) {
                        r1v = a.c.a1 * Dttt + a.c.b1 * (T(4.0) * cur.r1m1[r] - cur.r1m2[r]);
                        r2v = a.c.a2 * Dttt + a.c.b2 * (T(4.0) * cur.r2m1[r] - cur.r2m2[r]);
                        sumR = (r1v + r2v) * a.c.cdt2;
                    }
                    lap = a.c.kxy[reg] * (Dxx + Dyy) + a.c.kz[reg] * Dzz;
                    const T inc = d1 + lap + a.c.dlt[reg] * Dttt + a.c.bet[reg] * N12 + sumR;
                    pn = pc + inc;
                } else {
                    // W:203-234, reference expression order
                    Dxx = (T(-490.0) * pc + T(270.0) * sx1 - T(27.0) * sx2 + T(2.0) * sx3) / T(180.0);
                    Dyy = (T(-490.0) * pc + T(270.0) * sy1 - T(27.0) * sy2 + T(2.0) * sy3) / T(180.0);
                    Dzz = (T(-490.0) * pc + T(270.0) * sz1 - T(27.0) * sz2 + T(2.0) * sz3) / T(180.0);
                    Dttt = T(5.0) * pc - T(18.0) * pm1 + T(24.0) * pm2 - T(14.0) * pm3 + T(3.0) * pm4;
                    const T A = T(3.0) * pc - T(4.0) * pm1 + pm2;
                    N12 = (A * A) / T(4.0) + pc * (T(2.0) * pc - T(5.0) * pm1 + T(4.0) * pm2 - pm3);
                    T sumR = T(0);
                    if (reg) {
                        const T Dttt_2dt = Dttt / a.c.two_dt3;
                        r1v = (a.c.K1 * Dttt_2dt / a.c.den1) + ((a.c.tau1 * (T(4.0) * cur.r1m1[r] - cur.r1m2[r])) / a.c.den1b);
                        r2v = (a.c.K2 * Dttt_2dt / a.c.den2) + ((a.c.tau2 * (T(4.0) * cur.r2m1[r] - cur.r2m2[r])) / a.c.den2b);
                        sumR = (r1v + r2v) * a.c.cdt2;
                    }
                    lap = a.c.kxy[reg] * (Dxx + Dyy) + a.c.kz[reg] * Dzz;
                    pn = T(2.0) * pc - pm1 + a.c.kxy[reg] * (Dxx + Dyy) + a.c.kz[reg] * Dzz +
                         a.c.dlt[reg] * Dttt + a.c.bet[reg] * N12 + sumR;
                }
                if (W > 0) {
                    int dx = N - 1 - ix, dy = N - 1 - iyr[r], dz = a.g.Nz - 1 - z;
                    if (a.g.full_domain) { dx = min(dx, ix); dy = min(dy, iyr[r]); }
                    if (dx < W || dy < W || dz < W) {
                        T sp = T(0);
                        if (dx < W) { const T u = T(a.g.sponge_alpha) * T(W - dx); sp += u * u; }
                        if (dy < W) { const T u = T(a.g.sponge_alpha) * T(W - dy); sp += u * u; }
                        if (dz < W) { const T u = T(a.g.sponge_alpha) * T(W - dz); sp += u * u; }
                        pn = exp(-sp) * pn;
                    }
                }
                const uint32_t o = cidx[r] * ES;
                bstore<T>(pn, make_rsrc(a.pn + zo, pbytes), o);
                if (reg) {
                    bstore<T>(r1v, make_rsrc(a.r1 + zo, pbytes), o);
                    bstore<T>(r2v, make_rsrc(a.r2 + zo, pbytes), o);
                }
                if (pm_upd) {
                    // oldest level first, as one update per step would see them
                    const int mk = a.pmax_mask;
                    T m = cur.pmx[r];
                    bool up = false;
                    if ((mk & 32) && pm4 > m) { m = pm4; up = true; }
                    if ((mk & 16) && pm3 > m) { m = pm3; up = true; }
                    if ((mk & 8) && pm2 > m) { m = pm2; up = true; }
                    if ((mk & 4) && pm1 > m) { m = pm1; up = true; }
                    if ((mk & 2) && pc > m) { m = pc; up = true; }
                    if ((mk & 1) && pn > m) { m = pn; up = true; }
                    if (up) bstore<T>(m, make_rsrc(a.pmax + zo, pbytes), o);
                }
                if (a.focus && (long long)(zo + cidx[r]) == a.focus_idx) a.focus[step] = (double)pn;
                lastp[r] = pn;
                if (a.g.obj_on) {
                    const int dx = ix - a.g.ocx, dy = iyr[r] - a.g.ocy, dz = z - a.g.ocz;
                    if ((double)(dx * dx + dy * dy + dz * dz) <= a.g.or2) e0.add(pn * pn);
                    else e1.add(pn * pn);
                } else {
                    const T dd = pn - pc;
                    e0.add(pn * pn * T(a.g.escale[reg]));
                    e1.add(dd * dd - pn * lap);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
#pragma unroll
            for (int k = 0; k < 6; ++k) q[r][k] = q[r][k + 1];
        }
        if (PF == 1) {
            cur = nxt;
        } else if (PF == 2) {
            cur = nx1;
            nx1 = nxt;
        } else if (PF == 3) {
            cur = nx1;
            nx1 = nx2;
            nx2 = nxt;
        } else if (z + 1 < z1) {
            load_plane(z + 1, cur);
        }
    }
    // z-max faces (W:108-113), innermost first, from the new p at plane Nz-4
    if (a.fuse_z && zc == a.g.nzc - 1) {
        const int Nz = a.g.Nz;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (!act[r]) continue;
            const size_t c = (size_t)(Nz - 4) * plane + cidx[r];
            T u0 = T(0), u1 = T(0), u2 = T(0);
            if (a.boundary != BD_BC_ZERO) {
                const T n0 = a.p[c], n1 = a.p[c + plane], n2 = a.p[c + 2 * plane], n3 = a.p[c + 3 * plane];
                u0 = mur<T>(n0, lastp[r], n1, a.kz_t);
                u1 = mur<T>(n1, u0, n2, a.kz_t);
                u2 = mur<T>(n2, u1, n3, a.kz_t);
            }
            a.pn[c + plane] = u0;
            a.pn[c + 2 * plane] = u1;
            a.pn[c + 3 * plane] = u2;
            if (a.pmax) {
                if (u0 > a.pmax[c + plane]) a.pmax[c + plane] = u0;
                if (u1 > a.pmax[c + 2 * plane]) a.pmax[c + 2 * plane] = u1;
                if (u2 > a.pmax[c + 3 * plane]) a.pmax[c + 3 * plane] = u2;
            }
        }
    }
    // Fused x/y faces (W:88-135, the order of wv_xyfaces): x layers of the
    // inside rows, y layers of the inside columns, then the corners from the
    // x layers of the owner row, for every plane this workgroup wrote (its
    // march planes and, for the first / last z-chunk, the z-face planes).
    // The owners' new values were written by this workgroup before the barrier.
    if (a.fuse_xy) {
        __syncthreads();
        const bool zero = a.boundary == BD_BC_ZERO;
        const int lo = a.g.xlo;
        const int cx0 = a.g.xlo + bx * TX, cx1 = min(cx0 + TX, N);
        const int ry1 = min(ybase + ROWS, N);
        const int oxh = (N - 4 >= cx0 && N - 4 < cx1) ? 1 : 0;
        const int oxl = (lo && 3 >= cx0 && 3 < cx1) ? 1 : 0;
        const int oyh = (N - 4 >= ybase && N - 4 < ry1) ? 1 : 0;
        const int oyl = (lo && 3 >= ybase && 3 < ry1) ? 1 : 0;
        const int ra = max(ybase, lo), nr = max(0, min(ry1, N - 3) - ra);   // inside rows of the tile
        const int ca = max(cx0, lo), nc = max(0, min(cx1, N - 3) - ca);     // inside columns
        const int nxs = oxh + oxl, nys = oyh + oyl;
        const int nch = nxs * nr + nys * nc;             // chains of one plane
        const int ncor = nxs * nys;                      // corners of one plane
        const int nm = z1 - z0;
        const int nfirst = (a.fuse_z && zc == 0) ? 3 : 0;
        const int nlast = (a.fuse_z && zc == a.g.nzc - 1) ? 3 : 0;
        const int npl = nm + nfirst + nlast;
        auto plane_of = [&](int i) { return i < nm ? z0 + i : (i - nm < nfirst ? i - nm : a.g.Nz - 3 + (i - nm - nfirst)); };
        const T* __restrict__ P0 = a.p;
        T* Pn = a.pn;
        T* PM = a.pmax;
        const bool pm = PM != nullptr;
        // chains, KB per thread per batch: all loads of the batch, then the stores
        constexpr int KB = sizeof(T) == 4 ? 8 : 4;
        const int tch = npl * nch;
        for (int t0 = 0; t0 < tch; t0 += KB * NT) {
            int zz[KB], off[KB], d[KB];       // plane, in-plane offset of the owner, layer stride
            T v[KB], m[KB][4], q[KB][3];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int t = t0 + j * NT + (int)threadIdx.x;
                zz[j] = -1; off[j] = 0; d[j] = 0;
                v[j] = T(0);
                m[j][0] = m[j][1] = m[j][2] = m[j][3] = T(0);
                q[j][0] = q[j][1] = q[j][2] = T(0);
                if (t < tch) {
                    zz[j] = plane_of(t / nch);
                    int u = t % nch, ox, oy;
                    if (u < nxs * nr) {
                        const bool hi = oxh && u / nr == 0;
                        ox = hi ? N - 4 : 3; oy = ra + u % nr; d[j] = hi ? 1 : -1;
                    } else {
                        u -= nxs * nr;
                        const bool hi = oyh && u / nc == 0;
                        ox = ca + u % nc; oy = hi ? N - 4 : 3; d[j] = hi ? N : -N;
                    }
                    off[j] = oy * N + ox;
                    const T* p0 = P0 + (size_t)zz[j] * plane + off[j];
                    v[j] = Pn[(size_t)zz[j] * plane + off[j]];
#pragma unroll
                    for (int k = 0; k < 4; ++k) m[j][k] = p0[k * d[j]];
                    if (pm) {
                        const T* q0 = PM + (size_t)zz[j] * plane + off[j];
#pragma unroll
                        for (int k = 0; k < 3; ++k) q[j][k] = q0[(k + 1) * d[j]];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                if (zz[j] < 0) continue;
                const double kf = zz[j] < a.g.iz_tissue ? a.kxy_w : a.kxy_t;
                T w[3] = {T(0), T(0), T(0)};
                if (!zero) {
                    w[0] = mur<T>(m[j][0], v[j], m[j][1], kf);
                    w[1] = mur<T>(m[j][1], w[0], m[j][2], kf);
                    w[2] = mur<T>(m[j][2], w[1], m[j][3], kf);
                }
                T* o = Pn + (size_t)zz[j] * plane + off[j];
                T* om = pm ? PM + (size_t)zz[j] * plane + off[j] : nullptr;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    o[(k + 1) * d[j]] = w[k];
                    if (pm && w[k] > q[j][k]) om[(k + 1) * d[j]] = w[k];
                }
            }
        }
        // corners: the owner row's x layers (as its chain forms them), then
        // the y layers of the three x-face columns
        for (int t = threadIdx.x; t < npl * ncor; t += NT) {
            const int z = plane_of(t / ncor), u = t % ncor;
            const bool xhi = oxh && u / nys == 0, yhi = oyh && u % nys == 0;
            const int sx = xhi ? 1 : -1;
            const long long sy = yhi ? N : -N;
            const long long o = (long long)z * (long long)plane + (long long)(yhi ? N - 4 : 3) * N + (xhi ? N - 4 : 3);
            const double kf = z < a.g.iz_tissue ? a.kxy_w : a.kxy_t;
            T c[4][4], old[3][3];
            const T vo = Pn[o];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int k = 0; k < 4; ++k) c[j][k] = (j == 0 && k > 0) ? T(0) : P0[o + j * sx + k * sy];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) old[j][k] = pm ? PM[o + (j + 1) * sx + (k + 1) * sy] : T(0);
            T xv[3] = {T(0), T(0), T(0)};
            if (!zero) {
                xv[0] = mur<T>(c[0][0], vo, c[1][0], kf);
                xv[1] = mur<T>(c[1][0], xv[0], c[2][0], kf);
                xv[2] = mur<T>(c[2][0], xv[1], c[3][0], kf);
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T w[3] = {T(0), T(0), T(0)};
                if (!zero) {
                    w[0] = mur<T>(c[j + 1][0], xv[j], c[j + 1][1], kf);
                    w[1] = mur<T>(c[j + 1][1], w[0], c[j + 1][2], kf);
                    w[2] = mur<T>(c[j + 1][2], w[1], c[j + 1][3], kf);
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const long long i = o + (j + 1) * sx + (k + 1) * sy;
                    Pn[i] = w[k];
                    if (pm && w[k] > old[j][k]) PM[i] = w[k];
                }
            }
        }
    }
    if (a.epart) {
        const double s0 = block_sum<NT>(e0.value(), red);
        const double s1 = block_sum<NT>(e1.value(), red);
        if (threadIdx.x == 0) {
            double* o = a.epart + ((size_t)step * a.g.nblocks + blockIdx.x) * 2;
            o[0] = s0;
            o[1] = s1;
        }
    }
}

// Tile shapes the interior kernel is instantiated for.
struct WvVariant { int tx, ty, ry, pf; };
// pf: planes of loads in flight ahead of the plane being computed (0..3)
constexpr WvVariant kWvVariants[] = {
    {64, 4, 1, 0}, {64, 4, 1, 1}, {64, 4, 2, 1}, {64, 8, 1, 1},
    {64, 8, 2, 1}, {64, 4, 4, 1}, {64, 8, 2, 0}, {32, 8, 2, 1},
    {128, 4, 1, 0}, {64, 4, 2, 2}, {64, 4, 1, 2}, {64, 8, 1, 2},
    {128, 4, 1, 2}, {64, 4, 1, 3},
};
constexpr int kWvNumVariants = sizeof(kWvVariants) / sizeof(kWvVariants[0]);

}  // namespace bd
