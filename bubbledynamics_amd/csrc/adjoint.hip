// Discrete adjoint of the Westervelt step for phased-array gradients (gfx950).
//
// Forward model (the library's own step, full domain, boundary layers held at
// zero, sponge G, hard array source on plane iz=2):
//   E^n   = 2p^n - p^{n-1} + kL p^n + dlt D3^n + bet N^n + s (R1^n + R2^n)
//   p^{n+1} = G E^n (interior),  p^{n+1} = A_e sin(w t_{n+1} + phi_e) (source cells)
//   D3^n  = 5p^n - 18p^{n-1} + 24p^{n-2} - 14p^{n-3} + 3p^{n-4}
//   N^n   = A^2/4 + p^n B,  A = 3p^n - 4p^{n-1} + p^{n-2},  B = 2p^n - 5p^{n-1} + 4p^{n-2} - p^{n-3}
//   R_v^n = a_v D3^n + b_v (4 R_v^{n-1} - R_v^{n-2})
// Objective J = sum_{n=1..Nt} [ -wT sum_target (p^n)^2 + wO sum_outside (p^n)^2 ].
//
// Backward step n (lam^{n+1} complete -> lam^n), with w = G lam^{n+1} on interior cells:
//   rho_v^n = s w + 4 b_v rho_v^{n+1} - b_v rho_v^{n+2}
//   d^n = dlt w + a_1 rho_1^n + a_2 rho_2^n,   e^n = bet w
//   lam^n  = acc[n] + 2w + 5d + e (3A/2 + B + 2p^n) + (kL)^T w + dJ/dp^n
//   acc[n-1] += -w - 18d + e(-2A - 5p^n);  acc[n-2] += 24d + e(A/2 + 4p^n)
//   acc[n-3]  = 3d^{n+1} + (-14d - e p^n)
// (acc[n-3] starts with the 3 d^{n+1} that step n+1 contributes to it:
// d^{n+1} = dlt w^{n+2} + a_1 rho_1^{n+1} + a_2 rho_2^{n+1} is formed again
// here from w^{n+2} = G lam^{n+2}, the third level of the lam ring, and the
// rho^{n+1} this step reads anyway -- the same operations on the same values
// as step n+1, so the same bits -- instead of being stored by step n+1 and
// read back: 19 words per cell instead of 20.)
// At a source cell lam^n = (kL)^T w = kz (270 w(3) - 27 w(4) + 2 w(5)) / 180, and
// dJ/dA_e = sum_n sum_{c in e} lam^n_c sin(w t_n + phi_e), dJ/dphi_e = ... A_e cos(...).
// The forward levels a backward segment needs are recomputed from checkpoints
// taken every K steps (K >= n_steps keeps every level instead).
#include "common.h"
#include "wv_march.h"
#include "wv_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace bd {

template <typename T>
struct AdjArgs {
    const T* lam1; T* lam0; const T* lam2;     // lam^{n+1} (stencil), lam^n (out), lam^{n+2}
    const T* p0; const T* p1; const T* p2; const T* p3;
    const T* acc0; T* acc1; T* acc2; T* acc3;  // acc[n] (in), acc[n-1], acc[n-2] (in/out), acc[n-3] (out)
    const T* r1a; const T* r1b; T* r1o;
    const T* r2a; const T* r2b; T* r2o;
    T* lsrc;                          // lam^n on plane iz=2 (N*N), nullable
    T kxy, kz, dlt, bet, s, a1, b1, a2, b2;
    int N, Nz, sponge_w;
    double sponge_alpha;
    int ocx, ocy, ocz;
    double or2, wT, wO;
    int ntx, nty, nzc, zchunk, nblocks;
    int zb;                           // border columns of interior rows store their zeros
};

template <typename T>
__device__ __forceinline__ T sponge_g(int ix, int iy, int iz, int N, int Nz, int W, double alpha) {
    if (W <= 0) return T(1);
    int dx = min(N - 1 - ix, ix), dy = min(N - 1 - iy, iy), dz = Nz - 1 - iz;
    if (!(dx < W || dy < W || dz < W)) return T(1);
    T sp = T(0);
    if (dx < W) { const T u = T(alpha) * T(W - dx); sp += u * u; }
    if (dy < W) { const T u = T(alpha) * T(W - dy); sp += u * u; }
    if (dz < W) { const T u = T(alpha) * T(W - dz); sp += u * u; }
    return exp(-sp);
}

// Tile shapes of the backward kernel (TX x TY threads, one column each).
struct AdjTile { int ax, ay; };
constexpr AdjTile kAdjTiles[] = {{64, 4}, {64, 8}, {32, 8}, {128, 4}, {128, 8}};
constexpr int kAdjDefaultTile = 0, kAdjWideTile = 3;

// Loads of one z plane for one thread (issued a plane ahead of their use).
template <typename T>
struct AdjPlane {
    T wq;                       // w at (ix, iy, z+3): the z queue
    T hx, hyb, hyt;             // x halo and bottom / top y halo of plane z (edge threads only)
    T p0, p1, p2, p3;           // p^n .. p^{n-3}
    T c0, c1, c2;               // acc[n] .. acc[n-2]
    T w2;                       // w^{n+2} = G lam^{n+2}
    T r1a, r1b, r2a, r2b;       // rho_v^{n+1}, rho_v^{n+2}
};

// The adjoint field is stored as w = G lam (zero outside the interior), so
// the sponge factor is formed once per cell when lam^n is written.  Tiles
// start at x = 0, so the row segments a wavefront loads and stores are
// line-aligned; with zb the lanes of the border columns x < 3 and x > N-4 of
// an interior row store the zeros those cells hold (the buffers start
// zeroed and no other write reaches them), so the six stores cover whole
// lines instead of leaving partly written ones at both ends of every row
// (measured: 1 us of 32 at 128^3).
template <typename T, int kAX, int kAY, int PF>
__global__ void __launch_bounds__(kAX * kAY) adj_step_kernel(AdjArgs<T> a) {
    static_assert(PF == 1 || PF == 2, "planes of loads in flight: 1 or 2");
    static_assert(kAY >= 3 && kAX >= 6, "tile too small for the 3-cell halo");
    __shared__ T s[2][kAY + 6][kAX + 6];
    const int lb = xcd_remap((int)blockIdx.x, a.nblocks);
    const int zc = lb / (a.ntx * a.nty), t2 = lb % (a.ntx * a.nty);   // tile-major
    const int bx = t2 % a.ntx, by = t2 / a.ntx;
    const int tx = threadIdx.x % kAX, ty = threadIdx.x / kAX;
    const int N = a.N, Nz = a.Nz;
    const size_t plane = (size_t)N * N;
    const int ix = bx * kAX + tx, iy = 3 + by * kAY + ty;
    const bool inr = ix < N && iy < N;
    const bool act = ix >= 3 && ix <= N - 4 && iy <= N - 4;
    const bool zst = a.zb && inr && !act && iy <= N - 4;    // border column of an interior row
    const int z0 = 3 + zc * a.zchunk;
    const int z1 = min(z0 + a.zchunk, Nz - 3);
    constexpr uint32_t ES = sizeof(T);
    const uint32_t pbytes = (uint32_t)(plane * ES);
    const uint32_t o = (uint32_t)(iy * N + ix) * ES;
    // halo sources: x halo left (tx < 3) / right (tx >= kAX-3), y halo bottom / top
    const bool xL = tx < 3, xR = tx >= kAX - 3, yB = ty < 3, yT = ty >= kAY - 3;
    const int hxo = xL ? -3 : 3;                       // x offset of this thread's halo cell
    // Halo cells beyond the grid are read only by threads outside the
    // interior; their loads stay inside the field or return zero.  An x halo
    // left of x = 0 (first tile) is read by no interior lane and not loaded.
    const bool hxin = ix + hxo >= 0;
    auto load_plane = [&](int z, AdjPlane<T>& v) {
        const size_t zo = (size_t)z * plane;
        const rsrc_t rw = make_rsrc(a.lam1 + zo, (size_t)(Nz - z) * plane * ES);
        v.wq = inr ? bload<T>(rw, o, 3 * pbytes) : T(0);
        v.hx = (xL || xR) && hxin && iy < N ? bload<T>(rw, (uint32_t)((int)o + hxo * (int)ES)) : T(0);
        v.hyb = yB && ix < N ? bload<T>(rw, (uint32_t)((int)o - 3 * N * (int)ES)) : T(0);
        v.hyt = yT && ix < N ? bload<T>(rw, (uint32_t)((int)o + 3 * N * (int)ES)) : T(0);
        if (act) {
            v.p0 = bload<T>(make_rsrc(a.p0 + zo, pbytes), o);
            v.p1 = bload<T>(make_rsrc(a.p1 + zo, pbytes), o);
            v.p2 = bload<T>(make_rsrc(a.p2 + zo, pbytes), o);
            v.p3 = bload<T>(make_rsrc(a.p3 + zo, pbytes), o);
            v.c0 = bload<T>(make_rsrc(a.acc0 + zo, pbytes), o);
            v.c1 = bload<T>(make_rsrc(a.acc1 + zo, pbytes), o);
            v.c2 = bload<T>(make_rsrc(a.acc2 + zo, pbytes), o);
            v.w2 = bload<T>(make_rsrc(a.lam2 + zo, pbytes), o);
            v.r1a = bload<T>(make_rsrc(a.r1a + zo, pbytes), o);
            v.r1b = bload<T>(make_rsrc(a.r1b + zo, pbytes), o);
            v.r2a = bload<T>(make_rsrc(a.r2a + zo, pbytes), o);
            v.r2b = bload<T>(make_rsrc(a.r2b + zo, pbytes), o);
        }
    };
    T q[7];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = inr ? a.lam1[(size_t)(z0 - 3 + k) * plane + (size_t)iy * N + ix] : T(0);
    q[6] = T(0);
    if (zc == 0 && act && a.lsrc) {
        // source plane iz=2: only the z stencil reaches interior cells
        a.lsrc[(size_t)iy * N + ix] = a.kz * (T(270.0) * q[3] - T(27.0) * q[4] + T(2.0) * q[5]) / T(180.0);
    }
    // the loads of the next PF planes are in flight while plane z is computed
    AdjPlane<T> cur, nx1;
    load_plane(z0, cur);
    if (PF == 2 && z0 + 1 < z1) load_plane(z0 + 1, nx1);
    for (int z = z0; z < z1; ++z) {
        AdjPlane<T> nxt;
        if (z + PF < z1) load_plane(z + PF, nxt);
        const int b = (z - z0) & 1;
        q[6] = cur.wq;
        s[b][ty + 3][tx + 3] = q[3];
        if (xL || xR) s[b][ty + 3][tx + 3 + hxo] = cur.hx;
        if (yB) s[b][ty][tx + 3] = cur.hyb;
        if (yT) s[b][ty + 6][tx + 3] = cur.hyt;
        __syncthreads();
        if (act) {
            const size_t zo = (size_t)z * plane;
            const T w = q[3];
            const T* row = &s[b][ty + 3][tx + 3];
            constexpr T i180 = T(1.0 / 180.0);
            const T Lx = (T(-490.0) * w + T(270.0) * (row[1] + row[-1]) - T(27.0) * (row[2] + row[-2]) +
                          T(2.0) * (row[3] + row[-3])) * i180;
            const T Ly = (T(-490.0) * w + T(270.0) * (s[b][ty + 4][tx + 3] + s[b][ty + 2][tx + 3]) -
                          T(27.0) * (s[b][ty + 5][tx + 3] + s[b][ty + 1][tx + 3]) +
                          T(2.0) * (s[b][ty + 6][tx + 3] + s[b][ty][tx + 3])) * i180;
            const T Lz = (T(-490.0) * w + T(270.0) * (q[4] + q[2]) - T(27.0) * (q[5] + q[1]) +
                          T(2.0) * (q[6] + q[0])) * i180;
            const T ST = a.kxy * (Lx + Ly) + a.kz * Lz;
            const T pn = cur.p0, pm1 = cur.p1, pm2 = cur.p2, pm3 = cur.p3;
            const T A = T(3.0) * pn - T(4.0) * pm1 + pm2;
            const T B = T(2.0) * pn - T(5.0) * pm1 + T(4.0) * pm2 - pm3;
            const T rho1 = a.s * w + T(4.0) * a.b1 * cur.r1a - a.b1 * cur.r1b;
            const T rho2 = a.s * w + T(4.0) * a.b2 * cur.r2a - a.b2 * cur.r2b;
            const T d = a.dlt * w + a.a1 * rho1 + a.a2 * rho2;
            const T e = a.bet * w;
            const int dx = ix - a.ocx, dy = iy - a.ocy, dz = z - a.ocz;
            const bool tgt = (double)(dx * dx + dy * dy + dz * dz) <= a.or2;
            const T dj = tgt ? T(-2.0 * a.wT) * pn : T(2.0 * a.wO) * pn;
            const T lam = cur.c0 + T(2.0) * w + T(5.0) * d + e * (T(1.5) * A + B + T(2.0) * pn) + ST + dj;
            const T g = sponge_g<T>(ix, iy, z, N, Nz, a.sponge_w, a.sponge_alpha);
            bstore<T>(g * lam, make_rsrc(a.lam0 + zo, pbytes), o);
            bstore<T>(cur.c1 + (-w - T(18.0) * d + e * (T(-2.0) * A - T(5.0) * pn)), make_rsrc(a.acc1 + zo, pbytes), o);
            bstore<T>(cur.c2 + (T(24.0) * d + e * (T(0.5) * A + T(4.0) * pn)), make_rsrc(a.acc2 + zo, pbytes), o);
            const T d_next = a.dlt * cur.w2 + a.a1 * cur.r1a + a.a2 * cur.r2a;   // d^{n+1}
            bstore<T>(T(3.0) * d_next + (T(-14.0) * d - e * pn), make_rsrc(a.acc3 + zo, pbytes), o);
            bstore<T>(rho1, make_rsrc(a.r1o + zo, pbytes), o);
            bstore<T>(rho2, make_rsrc(a.r2o + zo, pbytes), o);
        } else if (zst) {
            const size_t zo = (size_t)z * plane;
            bstore<T>(T(0), make_rsrc(a.lam0 + zo, pbytes), o);
            bstore<T>(T(0), make_rsrc(a.acc1 + zo, pbytes), o);
            bstore<T>(T(0), make_rsrc(a.acc2 + zo, pbytes), o);
            bstore<T>(T(0), make_rsrc(a.acc3 + zo, pbytes), o);
            bstore<T>(T(0), make_rsrc(a.r1o + zo, pbytes), o);
            bstore<T>(T(0), make_rsrc(a.r2o + zo, pbytes), o);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = q[k + 1];
        if (PF == 1) {
            cur = nxt;
        } else {
            cur = nx1;
            nx1 = nxt;
        }
    }
}

// lam^{Nt} = dJ/dp^{Nt} on interior cells (nothing follows the last level),
// stored as w = G lam.
template <typename T>
__global__ void adj_init_kernel(T* lam, const T* p, int N, int Nz, int ocx, int ocy, int ocz, double or2, double wT,
                                double wO, int W, double alpha) {
    const size_t n = (size_t)N * N * Nz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ix = (int)(i % N), iy = (int)((i / N) % N), iz = (int)(i / ((size_t)N * N));
    T v = T(0);
    if (ix >= 3 && ix <= N - 4 && iy >= 3 && iy <= N - 4 && iz >= 3 && iz <= Nz - 4) {
        const int dx = ix - ocx, dy = iy - ocy, dz = iz - ocz;
        const bool tgt = (double)(dx * dx + dy * dy + dz * dz) <= or2;
        v = tgt ? T(-2.0 * wT) * p[i] : T(2.0 * wO) * p[i];
        v = sponge_g<T>(ix, iy, iz, N, Nz, W, alpha) * v;
    }
    lam[i] = v;
}

// Per-element gradient: one workgroup per element, FP64 accumulation over the
// element's source cells and all levels n = 1..Nt-1.
template <typename T>
__global__ void __launch_bounds__(256) adj_grad_kernel(const T* lsrc, const int* cells, const int* cell_off,
                                                       const double* amp, const double* phase, const double* times,
                                                       int N, int Nt, double w, double burst_cycles, double burst_t0,
                                                       double fr, double* gA, double* gP) {
    __shared__ double red[4];
    const int e = blockIdx.x;
    const int c0 = cell_off[e], c1 = cell_off[e + 1];
    const size_t plane = (size_t)N * N;
    double sa = 0.0, sp = 0.0;
    for (int n = 1 + (int)threadIdx.x; n < Nt; n += blockDim.x) {
        double S = 0.0;
        for (int c = c0; c < c1; ++c) S += (double)lsrc[(size_t)n * plane + cells[c]];
        double tb = times[n];
        bool on = true;
        if (burst_cycles > 0) {
            tb = times[n] - burst_t0;
            on = !(tb < 0.0 || tb * fr >= burst_cycles);
        }
        if (!on) continue;
        double sn, cs;
        sincos(w * tb + phase[e], &sn, &cs);
        sa += S * sn;
        sp += S * amp[e] * cs;
    }
    const double A = block_sum<256>(sa, red);
    const double P = block_sum<256>(sp, red);
    if (threadIdx.x == 0) { gA[e] = A; gP[e] = P; }
}

// J partial sums: a single workgroup, fixed order.
__global__ void __launch_bounds__(256) adj_sum_kernel(const double* epart, size_t n2, double* out) {
    __shared__ double red[4];
    double a = 0.0, b = 0.0;
    for (size_t i = threadIdx.x; i < n2; i += blockDim.x) { a += epart[2 * i]; b += epart[2 * i + 1]; }
    const double A = block_sum<256>(a, red);
    const double B = block_sum<256>(b, red);
    if (threadIdx.x == 0) { out[0] = A; out[1] = B; }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t bytes) { return hipMalloc(&p, bytes) == hipSuccess && hipMemset(p, 0, bytes) == hipSuccess; }
};

template <typename T>
static int adjoint_t(const bd_westervelt_params* prm, int Nxy, int Nz, int Nt, int K, const double* amp,
                     const double* phase, const bd_adjoint_objective* obj, double* J, double* gA, double* gP,
                     double* timings, hipStream_t st) {
    const int N = Nxy;
    const size_t plane = (size_t)N * N, NF = plane * Nz, FB = NF * sizeof(T);
    if (K <= 0 || K > Nt) K = Nt;
    const int nseg = (Nt + K - 1) / K;
    const int nb = wv_forward_nblocks(prm, sizeof(T) == 8 ? BD_F64 : BD_F32, Nxy, Nz);
    // forward workspace: segment store p^{a-4..a+K}, R rings, checkpoints
    DevBuf seg[2], rring, ckp, epart, jsum;
    if (!seg[0].alloc(FB * (size_t)(K + 5)) || !rring.alloc(FB * 6) || !epart.alloc(sizeof(double) * 2 * nb * (size_t)Nt) ||
        !jsum.alloc(sizeof(double) * 2))
        return BD_ERR_ALLOC;
    if (nseg > 1 && (!ckp.alloc(FB * 9 * (size_t)nseg) || !seg[1].alloc(FB * (size_t)(K + 5)))) return BD_ERR_ALLOC;
    // Segment j lives in store j&1: the recompute of segment j-1 fills one
    // store while the adjoint sweep of segment j reads the other.
    auto Sl = [&](int m, int a0) {   // p^m in the store of the segment starting at a0
        return (T*)seg[(a0 / K) & 1].p + (size_t)(m - (a0 - 4)) * NF;
    };
    T* Rr = (T*)rring.p;
    auto R1 = [&](int m) { return Rr + (size_t)(((m % 3) + 3) % 3) * NF; };
    auto R2 = [&](int m) { return Rr + (size_t)(3 + ((m % 3) + 3) % 3) * NF; };
    auto CK = [&](int j, int k) { return (T*)ckp.p + ((size_t)j * 9 + k) * NF; };
    std::vector<double> times((size_t)Nt + 1);
    {
        double t = 0.0;
        for (int k = 0; k <= Nt; ++k) { times[k] = t; t = t + prm->dt; }
    }
    WvObjective ob{obj->cx, obj->cy, obj->cz, obj->radius * obj->radius};
    auto step = [&](int m, int a0, double* ep, hipStream_t s_) -> int {
        WvStepPtrs q;
        q.pn = Sl(m + 1, a0); q.p = Sl(m, a0); q.pm1 = Sl(m - 1, a0); q.pm2 = Sl(m - 2, a0);
        q.pm3 = Sl(m - 3, a0); q.pm4 = Sl(m - 4, a0);
        q.r1 = R1(m); q.r1m1 = R1(m - 1); q.r1m2 = R1(m - 2);
        q.r2 = R2(m); q.r2m1 = R2(m - 1); q.r2m2 = R2(m - 2);
        return wv_forward_step(prm, sizeof(T) == 4 ? BD_F32 : BD_F64, Nxy, Nz, q, m, times[(size_t)m + 1], amp, phase,
                               ep, &ob, s_);
    };
    auto cp = [&](void* dst, const void* src, hipStream_t s_) {
        return hipMemcpyAsync(dst, src, FB, hipMemcpyDeviceToDevice, s_);
    };
    auto save_ck = [&](int j, int a0) -> int {  // state at step a0 (store of segment a0) into checkpoint j
        for (int k = 0; k < 5; ++k) BD_HIP_CHECK(cp(CK(j, k), Sl(a0 - k, a0), st));
        BD_HIP_CHECK(cp(CK(j, 5), R1(a0 - 1), st));
        BD_HIP_CHECK(cp(CK(j, 6), R1(a0 - 2), st));
        BD_HIP_CHECK(cp(CK(j, 7), R2(a0 - 1), st));
        BD_HIP_CHECK(cp(CK(j, 8), R2(a0 - 2), st));
        return BD_OK;
    };
    auto load_ck = [&](int j, int a0, hipStream_t s_) -> int {  // checkpoint j (state at a0) into its store and the R rings
        for (int k = 0; k < 5; ++k) BD_HIP_CHECK(cp(Sl(a0 - k, a0), CK(j, k), s_));
        BD_HIP_CHECK(cp(R1(a0 - 1), CK(j, 5), s_));
        BD_HIP_CHECK(cp(R1(a0 - 2), CK(j, 6), s_));
        BD_HIP_CHECK(cp(R2(a0 - 1), CK(j, 7), s_));
        BD_HIP_CHECK(cp(R2(a0 - 2), CK(j, 8), s_));
        return BD_OK;
    };
    auto now = [&]() {
        (void)hipStreamSynchronize(st);
        return std::chrono::steady_clock::now();
    };
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };

    // ---- forward pass: objective, checkpoints, last segment kept ----------
    auto t0 = now();
    for (int j = 0; j < nseg; ++j) {
        const int a0 = j * K, b = std::min(Nt, a0 + K);
        if (j > 0) {
            // the last five levels of segment j-1 head the store of segment j
            for (int k = 4; k >= 0; --k) BD_HIP_CHECK(cp(Sl(a0 - k, a0), Sl(a0 - k, a0 - K), st));
        }
        if (nseg > 1 && j > 0) { int rc = save_ck(j, a0); if (rc) return rc; }
        for (int m = a0; m < b; ++m) { int rc = step(m, a0, (double*)epart.p, st); if (rc) return rc; }
    }
    hipLaunchKernelGGL(adj_sum_kernel, dim3(1), dim3(256), 0, st, (const double*)epart.p, (size_t)nb * Nt,
                       (double*)jsum.p);
    BD_LAUNCH_CHECK();
    double js[2];
    BD_HIP_CHECK(hipMemcpyAsync(js, jsum.p, sizeof(js), hipMemcpyDeviceToHost, st));
    BD_HIP_CHECK(hipStreamSynchronize(st));
    *J = -obj->w_target * js[0] + obj->w_outside * js[1];
    auto t1 = now();
    if (timings) { timings[0] = ms(t0, t1); timings[1] = 0.0; timings[2] = 0.0; }
    if (!gA || !gP) return BD_OK;

    // ---- backward pass ----------------------------------------------------
    DevBuf lam, acc, rho, lsrc;
    if (!lam.alloc(FB * 3) || !acc.alloc(FB * 4) || !rho.alloc(FB * 6) || !lsrc.alloc(plane * sizeof(T) * (size_t)(Nt + 1)))
        return BD_ERR_ALLOC;
    T* L = (T*)lam.p;
    T* Ac = (T*)acc.p;
    T* Rh = (T*)rho.p;
    // lam ring of 3 levels (lam^{Nt+1} is the zero slot the sweep starts
    // with), accumulator ring of 4 (acc[m] is written at step m+3, updated at
    // m+2 and m+1 and read at m; the levels above Nt-4 are never written and
    // stay zero)
    auto Lv = [&](int m) { return L + (size_t)(((m % 3) + 3) % 3) * NF; };
    auto Av = [&](int m) { return Ac + (size_t)(((m % 4) + 4) % 4) * NF; };
    auto Rh1 = [&](int m) { return Rh + (size_t)(((m % 3) + 3) % 3) * NF; };
    auto Rh2 = [&](int m) { return Rh + (size_t)(3 + ((m % 3) + 3) % 3) * NF; };
    const WvCoef<T> cf = [&]() {
        WvCoef<T> c;
        const double dt = prm->dt;
        c.kxy[1] = (T)std::pow(prm->c_inf_tissue * dt / prm->dxy, 2);
        c.kz[1] = (T)std::pow(prm->c_inf_tissue * dt / prm->dz, 2);
        c.dlt[1] = (T)(prm->delta_tissue / (2.0 * dt * std::pow(prm->c_inf_tissue, 2)));
        c.bet[1] = (T)(2.0 * prm->beta_tissue / (prm->rho_tissue * std::pow(prm->c_inf_tissue, 2)));
        c.cdt2 = (T)std::pow(prm->c_inf_tissue * dt, 2);
        const double two_dt3 = 2.0 * std::pow(dt, 3), ct3 = std::pow(prm->c_inf_tissue, 3);
        c.a1 = (T)((2.0 * dt * 2.0 * prm->c1 * prm->tau1) / (two_dt3 * (2.0 * dt + 3.0 * prm->tau1) * ct3));
        c.b1 = (T)(prm->tau1 / (2.0 * dt + 3.0 * prm->tau1));
        c.a2 = (T)((2.0 * dt * 2.0 * prm->c2 * prm->tau2) / (two_dt3 * (2.0 * dt + 3.0 * prm->tau2) * ct3));
        c.b2 = (T)(prm->tau2 / (2.0 * dt + 3.0 * prm->tau2));
        return c;
    }();
    AdjArgs<T> A;
    A.kxy = cf.kxy[1]; A.kz = cf.kz[1]; A.dlt = cf.dlt[1]; A.bet = cf.bet[1]; A.s = cf.cdt2;
    A.a1 = cf.a1; A.b1 = cf.b1; A.a2 = cf.a2; A.b2 = cf.b2;
    A.N = N; A.Nz = Nz; A.sponge_w = prm->sponge_cells; A.sponge_alpha = prm->sponge_alpha;
    A.ocx = obj->cx; A.ocy = obj->cy; A.ocz = obj->cz; A.or2 = obj->radius * obj->radius;
    A.wT = obj->w_target; A.wO = obj->w_outside;
    // 64 or 128 columns wide, whichever leaves fewer idle lanes (BD_ADJ_TILE overrides)
    int tile = kAdjDefaultTile;
    {
        const int w64 = (N + 63) / 64 * 64 - N, w128 = (N + 127) / 128 * 128 - N;
        if (w128 <= w64) tile = kAdjWideTile;
    }
    if (const char* v = std::getenv("BD_ADJ_TILE")) { const int x = std::atoi(v); if (x >= 0 && x < 5) tile = x; }
    const int kAX = kAdjTiles[tile].ax, kAY = kAdjTiles[tile].ay;
    // border zeros on (BD_ADJ_ZB=0 leaves the border columns unwritten, for measurement)
    A.zb = 1;
    if (const char* v = std::getenv("BD_ADJ_ZB")) A.zb = std::atoi(v) ? 1 : 0;
    A.ntx = (N + kAX - 1) / kAX;
    A.nty = (N - 6 + kAY - 1) / kAY;
    {
        // target workgroup count (BD_ADJ_BLOCKS overrides it for the tile sweep)
        int target = 1024;
        if (const char* v = std::getenv("BD_ADJ_BLOCKS")) { const int x = std::atoi(v); if (x > 0) target = x; }
        const int nzin = Nz - 6;
        int nzc = std::max(1, target / (A.ntx * A.nty));
        // chunks of at least 4 planes (128^3, 128 x 4 tile: 4-plane chunks
        // 29.6 us per step, 8-plane 30.4, 16-plane 31.1; 6- and 12-plane
        // chunks leave a partly filled last round of workgroups, 34.6 / 36.6)
        int zchunk = std::max(std::min(4, nzin), (nzin + nzc - 1) / nzc);
        if (const char* v = std::getenv("BD_ADJ_ZCHUNK")) { const int x = std::atoi(v); if (x > 0) zchunk = std::min(x, nzin); }
        A.nzc = (nzin + zchunk - 1) / zchunk;
        A.zchunk = zchunk;
        A.nblocks = A.ntx * A.nty * A.nzc;
    }
    // recompute stream (default priority) and the events that order the two streams
    struct Side {
        hipStream_t s = nullptr;
        std::vector<hipEvent_t> ev;
        ~Side() {
            for (auto& e : ev) if (e) (void)hipEventDestroy(e);
            if (s) (void)hipStreamDestroy(s);
        }
    } side;
    BD_HIP_CHECK(hipStreamCreateWithFlags(&side.s, hipStreamNonBlocking));
    side.ev.assign(4 + 2 * (size_t)nseg, nullptr);
    for (auto& e : side.ev) BD_HIP_CHECK(hipEventCreate(&e));
    hipEvent_t* recomputed = side.ev.data();       // [store]: its recompute finished
    hipEvent_t* swept = side.ev.data() + 2;        // [store]: the adjoint sweep over it finished
    hipEvent_t* sweep_t = side.ev.data() + 4;      // [2j], [2j+1]: sweep of segment j starts / ends
    // recompute p^{a0+1..b} of segment j from its checkpoint (segment 0 from rest) on the side stream
    auto recompute = [&](int j) -> int {
        const int a0 = j * K, b = std::min(Nt, a0 + K), sl = j & 1;
        hipStream_t s2 = side.s;
        if (j + 2 < nseg) BD_HIP_CHECK(hipStreamWaitEvent(s2, swept[sl], 0));   // store last read by segment j+2
        if (j == 0) {
            BD_HIP_CHECK(hipMemsetAsync(Sl(-4, 0), 0, FB * 5, s2));
            BD_HIP_CHECK(hipMemsetAsync(Rr, 0, FB * 6, s2));
        } else {
            int rc = load_ck(j, a0, s2);
            if (rc) return rc;
        }
        for (int m = a0; m < b; ++m) { int rc = step(m, a0, nullptr, s2); if (rc) return rc; }
        BD_HIP_CHECK(hipEventRecord(recomputed[sl], s2));
        return BD_OK;
    };
    // planes of loads in flight (BD_ADJ_PF = 1 or 2, default 2)
    int pf = 2;
    if (const char* v = std::getenv("BD_ADJ_PF")) pf = std::atoi(v) == 1 ? 1 : 2;
    auto launch_adj = [&](hipStream_t s_) -> int {
        if (pf == 1) {
            switch (tile) {
                case 0: hipLaunchKernelGGL((adj_step_kernel<T, 64, 4, 1>), dim3(A.nblocks), dim3(256), 0, s_, A); break;
                case 1: hipLaunchKernelGGL((adj_step_kernel<T, 64, 8, 1>), dim3(A.nblocks), dim3(512), 0, s_, A); break;
                case 2: hipLaunchKernelGGL((adj_step_kernel<T, 32, 8, 1>), dim3(A.nblocks), dim3(256), 0, s_, A); break;
                case 3: hipLaunchKernelGGL((adj_step_kernel<T, 128, 4, 1>), dim3(A.nblocks), dim3(512), 0, s_, A); break;
                default: hipLaunchKernelGGL((adj_step_kernel<T, 128, 8, 1>), dim3(A.nblocks), dim3(1024), 0, s_, A); break;
            }
        } else {
            switch (tile) {
                case 0: hipLaunchKernelGGL((adj_step_kernel<T, 64, 4, 2>), dim3(A.nblocks), dim3(256), 0, s_, A); break;
                case 1: hipLaunchKernelGGL((adj_step_kernel<T, 64, 8, 2>), dim3(A.nblocks), dim3(512), 0, s_, A); break;
                case 2: hipLaunchKernelGGL((adj_step_kernel<T, 32, 8, 2>), dim3(A.nblocks), dim3(256), 0, s_, A); break;
                case 3: hipLaunchKernelGGL((adj_step_kernel<T, 128, 4, 2>), dim3(A.nblocks), dim3(512), 0, s_, A); break;
                default:
                    // 1024 threads leave 128 VGPRs: the FP64 two-plane pipeline would spill
                    if constexpr (sizeof(T) == 8)
                        hipLaunchKernelGGL((adj_step_kernel<T, 128, 8, 1>), dim3(A.nblocks), dim3(1024), 0, s_, A);
                    else
                        hipLaunchKernelGGL((adj_step_kernel<T, 128, 8, 2>), dim3(A.nblocks), dim3(1024), 0, s_, A);
                    break;
            }
        }
        BD_LAUNCH_CHECK();
        return BD_OK;
    };
    double t_sweep = 0.0;
    const auto p0 = std::chrono::steady_clock::now();
    {
        // lam^{Nt}
        const int last_a0 = (nseg - 1) * K;
        hipLaunchKernelGGL((adj_init_kernel<T>), dim3((unsigned)((NF + 255) / 256)), dim3(256), 0, st, Lv(Nt),
                           (const T*)Sl(Nt, last_a0), N, Nz, obj->cx, obj->cy, obj->cz, A.or2, obj->w_target,
                           obj->w_outside, prm->sponge_cells, prm->sponge_alpha);
        BD_LAUNCH_CHECK();
    }
    // Segments nseg-1 and nseg-2 are still in the two stores after the
    // forward pass; every earlier segment is recomputed into the store of
    // segment j+2 once the sweep over it is done, overlapping the sweep of j+1.
    if (nseg > 1) BD_HIP_CHECK(hipEventRecord(recomputed[(nseg - 2) & 1], st));
    for (int j = nseg - 1; j >= 0; --j) {
        const int a0 = j * K, b = std::min(Nt, a0 + K), sl = j & 1;
        if (j != nseg - 1) BD_HIP_CHECK(hipStreamWaitEvent(st, recomputed[sl], 0));
        BD_HIP_CHECK(hipEventRecord(sweep_t[2 * j], st));
        for (int n = b - 1; n >= std::max(a0, 1); --n) {
            A.lam1 = Lv(n + 1); A.lam0 = Lv(n); A.lam2 = Lv(n + 2);
            A.p0 = Sl(n, a0); A.p1 = Sl(n - 1, a0); A.p2 = Sl(n - 2, a0); A.p3 = Sl(n - 3, a0);
            A.acc0 = Av(n); A.acc1 = Av(n - 1); A.acc2 = Av(n - 2); A.acc3 = Av(n - 3);
            A.r1a = Rh1(n + 1); A.r1b = Rh1(n + 2); A.r1o = Rh1(n);
            A.r2a = Rh2(n + 1); A.r2b = Rh2(n + 2); A.r2o = Rh2(n);
            A.lsrc = (T*)lsrc.p + (size_t)n * plane;
            int rc = launch_adj(st);
            if (rc) return rc;
        }
        BD_HIP_CHECK(hipEventRecord(sweep_t[2 * j + 1], st));
        BD_HIP_CHECK(hipEventRecord(swept[sl], st));
        // segment j-2 reuses this store: its recompute starts once this sweep is done
        if (j - 2 >= 0) { int rc = recompute(j - 2); if (rc) return rc; }
    }
    BD_HIP_CHECK(hipStreamSynchronize(side.s));
    BD_HIP_CHECK(hipStreamSynchronize(st));
    const double t_phase = ms(p0, std::chrono::steady_clock::now());
    for (int j = 0; j < nseg; ++j) {
        float e = 0.f;
        BD_HIP_CHECK(hipEventElapsedTime(&e, sweep_t[2 * j], sweep_t[2 * j + 1]));
        t_sweep += e;
    }
    // element cell lists (host), then the gradient reduction
    const int ne = prm->arr_nx * prm->arr_ny;
    std::vector<std::vector<int>> lists(ne);
    for (int iy = 3; iy <= N - 4; ++iy)
        for (int ix = 3; ix <= N - 4; ++ix) {
            const double X = prm->xy_min + ix * prm->dxy, Y = prm->xy_min + iy * prm->dxy;
            const double fx = std::floor((X - prm->arr_x0) / prm->arr_pitch);
            const double fy = std::floor((Y - prm->arr_y0) / prm->arr_pitch);
            if (fx < 0 || fy < 0 || fx >= prm->arr_nx || fy >= prm->arr_ny) continue;
            lists[(int)fy * prm->arr_nx + (int)fx].push_back(iy * N + ix);
        }
    std::vector<int> cells, off(ne + 1, 0);
    for (int e = 0; e < ne; ++e) { off[e + 1] = off[e] + (int)lists[e].size(); cells.insert(cells.end(), lists[e].begin(), lists[e].end()); }
    DevBuf dcells, doff, dtimes;
    if (!dcells.alloc(sizeof(int) * std::max<size_t>(1, cells.size())) || !doff.alloc(sizeof(int) * off.size()) ||
        !dtimes.alloc(sizeof(double) * times.size()))
        return BD_ERR_ALLOC;
    BD_HIP_CHECK(hipMemcpyAsync(dcells.p, cells.data(), sizeof(int) * cells.size(), hipMemcpyHostToDevice, st));
    BD_HIP_CHECK(hipMemcpyAsync(doff.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, st));
    BD_HIP_CHECK(hipMemcpyAsync(dtimes.p, times.data(), sizeof(double) * times.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((adj_grad_kernel<T>), dim3(ne), dim3(256), 0, st, (const T*)lsrc.p, (const int*)dcells.p,
                       (const int*)doff.p, amp, phase, (const double*)dtimes.p, N, Nt, 2.0 * M_PI * prm->fr,
                       prm->burst_cycles, prm->burst_t0, prm->fr, gA, gP);
    BD_LAUNCH_CHECK();
    BD_HIP_CHECK(hipStreamSynchronize(st));
    if (timings) { timings[1] = std::max(0.0, t_phase - t_sweep); timings[2] = t_sweep; }
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_westervelt_adjoint(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz, int n_steps,
                                     int checkpoint_every, const double* elem_amp, const double* elem_phase,
                                     const bd_adjoint_objective* obj, double* J, double* grad_amp,
                                     double* grad_phase, double* timings_ms, void* stream) {
    int rc = wv_check_params(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!obj || !J || !elem_amp || !elem_phase || n_steps < 1) return BD_ERR_INVALID;
    if (!prm->full_domain || prm->boundary != BD_BC_ZERO || prm->source != BD_SRC_ARRAY) return BD_ERR_UNSUPPORTED;
    // one medium throughout (every plane in the relaxing branch)
    if (!(prm->z_min >= prm->water_tissue)) return BD_ERR_UNSUPPORTED;
    if (prm->arr_nx <= 0 || prm->arr_ny <= 0 || !(prm->arr_pitch > 0)) return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    if (dtype == BD_F32)
        return adjoint_t<float>(prm, Nxy, Nz, n_steps, checkpoint_every, elem_amp, elem_phase, obj, J, grad_amp,
                                grad_phase, timings_ms, st);
    return adjoint_t<double>(prm, Nxy, Nz, n_steps, checkpoint_every, elem_amp, elem_phase, obj, J, grad_amp,
                             grad_phase, timings_ms, st);
}
