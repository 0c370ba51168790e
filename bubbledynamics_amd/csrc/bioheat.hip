// Heat conduction with a moving acoustic source and CEM43 dose on gfx950.
//
// Reference: Westervelt_Temperature_SquareSpiral.cu (T:line).
//
// Kernel (bh_march): a workgroup owns a 64 x 4 tile of columns and a chunk of
// z planes; each thread marches z with a 3-deep register queue of its own
// column, the plane with a one-cell x/y halo is staged in LDS (double
// buffered, one barrier per plane), and Q (shifted), TD and the optional
// Kahan compensation are streamed once; every load of plane z+1 is issued
// before plane z is computed.  Fields may hold T - T_off so
// that FP32 resolves small temperature rises.
#include "common.h"
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

namespace bd {

template <typename T>
struct BhArgs {
    const T* src; T* dst; const T* Q; T* TD; T* TDc;
    int Nxy, Nz, iz_tissue, shift_x, shift_y;
    int ntx, nty, nzc, zchunk, nblocks;
    T kxy[2], kz[2], dq[2];            // [0] water, [1] tissue (T:40-45)
    T keep[2];                         // 1 - 4 kxy - 2 kz          (T:85-96)
    T perf, theta_art;                 // Pennes: perf * (T - T_art), tissue only
    double T_off;                      // stored value + T_off = T [K]
    double dt_T;
    T dose_off;                        // (T_off - 273.15) - 43, FP32 path
    int reference_order;               // FP64 drop-in: expression order of T:85-102
};

constexpr int kBX = 64, kBY = 4;

// Coefficients of one medium for one substep, selected once per plane.
// Perfusion acts in tissue only; elsewhere its coefficient is zero, and
// Tn - 0 * (Tc - theta) is Tn for every finite temperature.
template <typename T>
struct BhCoef { T keep, kxy, kz, dq, perf, theta; };

template <typename T>
__device__ __forceinline__ BhCoef<T> bh_coef(const BhArgs<T>& a, int reg) {
    return {a.keep[reg], a.kxy[reg], a.kz[reg], a.dq[reg], reg ? a.perf : T(0), a.theta_art};
}

// One cell of one substep (T:83-97) in the library's order: conduction,
// source, Pennes perfusion.
template <typename T>
__device__ __forceinline__ T bh_update_c(const BhCoef<T>& c, T Tc, T l, T r, T d, T u, T zp, T zm, T q) {
    const T Tn = Tc * c.keep + c.kxy * ((l + r) + (u + d)) + c.kz * (zp + zm) + c.dq * q;
    return Tn - c.perf * (Tc - c.theta);
}

// One cell of one substep: the reference's expression order (T:85-96) for
// the FP64 drop-in, else bh_update_c.
template <typename T>
__device__ __forceinline__ T bh_update(const BhArgs<T>& a, int reg, T Tc, T l, T r, T d, T u, T zp, T zm, T q) {
    if (!a.reference_order) return bh_update_c(bh_coef(a, reg), Tc, l, r, d, u, zp, zm, q);
    T Tn = Tc * (T(1.0) - T(4.0) * a.kxy[reg] - T(2.0) * a.kz[reg]) + a.kxy[reg] * (l + r + u + d) +
           a.kz[reg] * (zp + zm) + a.dq[reg] * q;
    if (reg && a.perf != T(0)) Tn = Tn - a.perf * (Tc - a.theta_art);
    return Tn;
}

// CEM43 dose constants: stored value + T_off = T [K]; dose_off = (T_off -
// 273.15) - 43 for the FP32 exponent.
template <typename T>
struct BhDoseC { T dose_off; double T_off, dt; };

// CEM43 dose increment of one substep at the new temperature (T:99-102).
template <typename T>
__device__ __forceinline__ T bh_dose_c(const BhDoseC<T>& c, T Tn) {
    if constexpr (std::is_same<T, float>::value) {
        // T_c - 43; the hardware exp2 equals exp2f wherever the result is a
        // normal number (T_c above -20 C)
        const float x = Tn + c.dose_off;
        return __builtin_amdgcn_exp2f(x < 0.0f ? 2.0f * x : x) * (float)c.dt;
    } else {
        // T:99-102 writes pow(4, T_c - 43) and pow(2, T_c - 43); they are
        // evaluated as exp2(2 (T_c - 43)) and exp2(T_c - 43), the same values
        // (2x is exact) up to the last bit of the library function, at a
        // fraction of pow's cost (FP64 pair kernel: 66 -> 149 G cell-substeps/s)
        const double T_c = (Tn + c.T_off) - 273.15;   // T:88, T:96
        return (T_c < 43.0) ? exp2(2.0 * (T_c - 43.0)) * c.dt : exp2(T_c - 43.0) * c.dt;
    }
}

template <typename T>
__device__ __forceinline__ T bh_dose(const BhArgs<T>& a, T Tn) {
    return bh_dose_c(BhDoseC<T>{a.dose_off, a.T_off, a.dt_T}, Tn);
}

// TD += add, Kahan-compensated when the compensation field is in use.
template <typename T>
__device__ __forceinline__ void bh_accumulate(T& td, T& tdc, T add, bool kahan) {
    if (!kahan) {
        td = td + add;
    } else {
        const T y = add - tdc, tt = td + y;
        tdc = (tt - td) - y;
        td = tt;
    }
}

// Loads of one plane for one thread, issued a plane ahead of their use.
template <typename T>
struct BhPlane { T zp, hx, hy, q, td, tdc; };

template <typename T>
__global__ void __launch_bounds__(kBX * kBY) bh_march(BhArgs<T> a) {
    __shared__ T s[2][kBY + 2][kBX + 2];
    const int lb = xcd_remap((int)blockIdx.x, a.nblocks);
    const int zc = lb / (a.ntx * a.nty), t2 = lb % (a.ntx * a.nty);   // tile-major
    const int bx = t2 % a.ntx, by = t2 / a.ntx;
    const int tx = (int)threadIdx.x % kBX, ty = (int)threadIdx.x / kBX;
    const int ix = 1 + bx * kBX + tx;
    const int iy = 1 + by * kBY + ty;
    const int N = a.Nxy;
    const bool inr = ix < N && iy < N;              // a grid cell (loads)
    const bool act = ix <= N - 2 && iy <= N - 2;    // an interior cell (update)
    const size_t plane = (size_t)N * N;
    const uint32_t ES = sizeof(T);
    const uint32_t pbytes = (uint32_t)(plane * ES);
    const int z0 = 1 + zc * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.Nz - 1);
    const uint32_t c = inr ? (uint32_t)iy * (uint32_t)N + (uint32_t)ix : 0u;
    const uint32_t o = c * ES;
    const int ixa = ix - a.shift_x, iya = iy - a.shift_y;
    const bool q_in = act && ixa >= 0 && ixa < N && iya >= 0 && iya < N;   // T:76
    const uint32_t cq = q_in ? ((uint32_t)iya * (uint32_t)N + (uint32_t)ixa) * ES : 0xFFFFFFFFu;
    // x / y halo cells of the tile (edge lanes / rows); a cell past the grid
    // is read only by threads outside the interior
    const int hxd = tx == 0 ? -1 : (tx == kBX - 1 ? 1 : 0);
    const int hyd = ty == 0 ? -1 : (ty == kBY - 1 ? 1 : 0);
    const bool hx_on = hxd != 0 && iy < N && ix + hxd < N;
    const bool hy_on = hyd != 0 && ix < N && iy + hyd < N;
    const uint32_t ohx = hx_on ? (uint32_t)((int)c + hxd) * ES : 0xFFFFFFFFu;
    const uint32_t ohy = hy_on ? (uint32_t)((int)c + hyd * N) * ES : 0xFFFFFFFFu;
    auto load = [&](int z, BhPlane<T>& v) {
        const size_t zo = (size_t)z * plane;
        const rsrc_t rs = make_rsrc(a.src + zo, 2 * (size_t)pbytes);
        v.zp = inr ? bload<T>(rs, o, pbytes) : T(0);
        v.hx = bload<T>(rs, ohx);                  // out of range offset: 0
        v.hy = bload<T>(rs, ohy);
        // out-of-range offset (shift outside the grid) reads 0, as T:75-77
        v.q = bload<T>(make_rsrc(a.Q + zo, pbytes), cq);
        v.td = act ? bload<T>(make_rsrc(a.TD + zo, pbytes), o) : T(0);
        v.tdc = (act && a.TDc) ? bload<T>(make_rsrc(a.TDc + zo, pbytes), o) : T(0);
    };

    T zm = inr ? a.src[(size_t)(z0 - 1) * plane + c] : T(0);
    T zc0 = inr ? a.src[(size_t)z0 * plane + c] : T(0);
    BhPlane<T> cur, nxt;
    load(z0, cur);
    for (int z = z0; z < z1; ++z) {
        if (z + 1 < z1) load(z + 1, nxt);
        const int b = (z - z0) & 1;
        s[b][ty + 1][tx + 1] = zc0;
        if (hxd) s[b][ty + 1][tx + 1 + hxd] = cur.hx;
        if (hyd) s[b][ty + 1 + hyd][tx + 1] = cur.hy;
        __syncthreads();
        if (act) {
            const size_t zo = (size_t)z * plane;
            const int reg = z >= a.iz_tissue ? 1 : 0;
            const T Tn = bh_update(a, reg, zc0, s[b][ty + 1][tx], s[b][ty + 1][tx + 2], s[b][ty][tx + 1],
                                   s[b][ty + 2][tx + 1], cur.zp, zm, cur.q);
            bstore<T>(Tn, make_rsrc(a.dst + zo, pbytes), o);
            T td = cur.td, tdc = cur.tdc;
            bh_accumulate(td, tdc, bh_dose(a, Tn), a.TDc != nullptr);
            if (a.TDc) bstore<T>(tdc, make_rsrc(a.TDc + zo, pbytes), o);
            bstore<T>(td, make_rsrc(a.TD + zo, pbytes), o);
        }
        zm = zc0;
        zc0 = cur.zp;
        cur = nxt;
    }
}

// Two substeps in one launch (temporal blocking).  The workgroup's BX x
// TY*RY region computes substep 1 on every region cell and substep 2 one
// plane behind on the inner (BX-2) x (TY*RY-2) cells, whose neighbours are
// all region cells; neighbouring regions overlap by two columns / rows.
// z: the workgroup owns output planes [z0, z1) and runs substep 1 on planes
// z0-1 .. z1.  Substep-1 values live in registers (own column, 3 planes) and
// in an LDS plane (x/y neighbours); the dose of both substeps is added before
// TD is stored once.  The per-cell arithmetic is bh_update_c / bh_dose_c /
// bh_accumulate in the order of two single substeps, so the result is
// bit-identical to them.  Substep-1 values of grid-boundary cells (never
// updated) are read from `dst`, which the first of two single substeps
// would have left there; the result goes to `dst`.
// Few instructions per cell: loads and stores of cells a lane does not own
// take an out-of-range buffer offset (loads return 0, stores are dropped),
// except the boundary and halo loads, which a wavefront without such a lane
// skips; every lane evaluates both updates and keeps what it owns; both
// media's coefficients come from LDS; the compensated dose is a template
// choice.  Every load of plane zs+1 is issued before plane zs is computed.
template <typename T, int BX, int TY, int RY, bool KAHAN>
__global__ void __launch_bounds__(BX * TY) bh_march2(BhArgs<T> a) {
    constexpr int ROWS = TY * RY;
    // the y-halo rows below and above the region come from the first and last
    // thread row, and a wavefront holds one thread row (uniform halo branches)
    static_assert(TY >= 2 && RY >= 1 && BX % 64 == 0, "two thread rows of whole wavefronts");
    constexpr uint32_t OOB = 0xFFFFFFFFu;
    constexpr bool kahan = KAHAN;
    __shared__ T sa[2][ROWS + 2][BX + 2];    // T^n plane zs with a one-cell halo
    __shared__ T sb[2][ROWS + 2][BX + 2];    // substep-1 plane (its ring is never read into a kept value)
    const int lb = xcd_remap((int)blockIdx.x, a.nblocks);
    const int zc = lb / (a.ntx * a.nty), t2 = lb % (a.ntx * a.nty);   // tile-major
    const int bx = t2 % a.ntx, by = t2 / a.ntx;
    const int tx = (int)threadIdx.x % BX, ty = (int)threadIdx.x / BX;   // ty < TY
    const int N = a.Nxy;
    const size_t plane = (size_t)N * N;
    const uint32_t ES = sizeof(T);
    const uint32_t pbytes = (uint32_t)(plane * ES);
    const int gx = bx * (BX - 2) + tx;
    const int gy0 = by * (ROWS - 2);
    const int z0 = 1 + zc * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.Nz - 1);
    const BhDoseC<T> dc{a.dose_off, a.T_off, a.dt_T};
    // both media's coefficients in LDS: a plane's medium is read there (one
    // broadcast read) instead of being selected into registers on every lane
    __shared__ T scoef[2][6];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const BhCoef<T> c = bh_coef(a, m);
            scoef[m][0] = c.keep; scoef[m][1] = c.kxy; scoef[m][2] = c.kz;
            scoef[m][3] = c.dq; scoef[m][4] = c.perf; scoef[m][5] = c.theta;
        }
    }
    auto coef = [&](int m) {
        return BhCoef<T>{scoef[m][0], scoef[m][1], scoef[m][2], scoef[m][3], scoef[m][4], scoef[m][5]};
    };

    // per row: offsets of the cell as a grid cell (o_in), as a grid cell that
    // substep 1 leaves unchanged (o_edge), as an output cell (o_out), of the
    // shifted source (cq) and of the x halo (ohx); OOB where the row's cell
    // is not one
    bool itr[RY];
    uint32_t o_in[RY], o_edge[RY], o_out[RY], cq[RY], ohx[RY];
    const int hxd = tx == 0 ? -1 : (tx == BX - 1 ? 1 : 0);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int ly = ty * RY + r;
        const int gy = gy0 + ly;
        const bool ing = gx < N && gy < N;
        itr[r] = gx >= 1 && gx <= N - 2 && gy >= 1 && gy <= N - 2;
        const bool outc = itr[r] && tx >= 1 && tx <= BX - 2 && ly >= 1 && ly <= ROWS - 2;
        const uint32_t o = ((uint32_t)gy * (uint32_t)N + (uint32_t)gx) * ES;
        o_in[r] = ing ? o : OOB;
        o_edge[r] = ing && !itr[r] ? o : OOB;
        o_out[r] = outc ? o : OOB;
        const int ixa = gx - a.shift_x, iya = gy - a.shift_y;
        const bool q_in = itr[r] && ixa >= 0 && ixa < N && iya >= 0 && iya < N;   // T:76
        cq[r] = q_in ? ((uint32_t)iya * (uint32_t)N + (uint32_t)ixa) * ES : OOB;
        const int hx = gx + hxd;
        ohx[r] = (hxd != 0 && gy < N && hx >= 0 && hx < N) ? ((uint32_t)gy * (uint32_t)N + (uint32_t)hx) * ES : OOB;
    }
    // halo rows of the region: row -1 (thread row 0) and row ROWS (last thread row)
    const int hyd = ty == 0 ? -1 : (ty == TY - 1 ? 1 : 0);
    const int hyrow = hyd < 0 ? gy0 - 1 : gy0 + ROWS;
    const uint32_t ohy = (hyd != 0 && gx < N && hyrow >= 0 && hyrow < N)
                             ? ((uint32_t)hyrow * (uint32_t)N + (uint32_t)gx) * ES : OOB;

    struct P2 { T zp[RY], q[RY], td[RY], tdc[RY], bnd[RY], hx[RY], hy; };
    // loads of step-1 plane zs (zs in [0, Nz-1]); zp is plane zs+1 (0 past the last plane)
    auto load = [&](int zs, P2& v) {
        const size_t zo = (size_t)zs * plane;
        const rsrc_t rs = make_rsrc(a.src + zo, pbytes);
        const rsrc_t rn = make_rsrc(a.src + zo + plane, zs + 1 < a.Nz ? pbytes : 0u);
        const rsrc_t rq = make_rsrc(a.Q + zo, pbytes);
        const rsrc_t rb = make_rsrc(a.dst + zo, pbytes);
        const rsrc_t rd = make_rsrc(a.TD + zo, pbytes);
        const bool bplane = zs == 0 || zs == a.Nz - 1;
        const bool outp = zs >= z0 && zs < z1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const uint32_t od = outp ? o_out[r] : OOB;
            const uint32_t ob = bplane ? o_in[r] : o_edge[r];
            v.zp[r] = bload<T>(rn, o_in[r]);
            v.q[r] = bload<T>(rq, cq[r]);
            // boundary and halo loads concern few lanes: a wavefront without
            // such a lane skips them
            v.bnd[r] = ob != OOB ? bload<T>(rb, ob) : T(0);
            v.td[r] = bload<T>(rd, od);
            v.hx[r] = ohx[r] != OOB ? bload<T>(rs, ohx[r]) : T(0);
        }
        if constexpr (KAHAN) {
            const rsrc_t rc = make_rsrc(a.TDc + zo, pbytes);
#pragma unroll
            for (int r = 0; r < RY; ++r) v.tdc[r] = bload<T>(rc, outp ? o_out[r] : OOB);
        } else {
#pragma unroll
            for (int r = 0; r < RY; ++r) v.tdc[r] = T(0);
        }
        v.hy = ohy != OOB ? bload<T>(rs, ohy) : T(0);
    };

    const int zb = z0 - 1, ze = z1;           // step-1 planes
    T tm[RY], tc[RY], t1m[RY], t1c[RY], qp[RY], hd[RY], hdc[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const rsrc_t rs = make_rsrc(a.src + (size_t)zb * plane, pbytes);
        const rsrc_t rm = make_rsrc(a.src + (size_t)(zb > 0 ? zb - 1 : 0) * plane, zb >= 1 ? pbytes : 0u);
        tm[r] = bload<T>(rm, o_in[r]);
        tc[r] = bload<T>(rs, o_in[r]);
        t1m[r] = t1c[r] = qp[r] = hd[r] = hdc[r] = T(0);
    }

    // one plane: cur holds the loads of plane zs; the loads of plane zs+1 go to nxt
    auto body = [&](int zs, P2& cur, P2& nxt) {
        if (zs + 1 <= ze) load(zs + 1, nxt);
        const int b = (zs - zb) & 1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int lp = ty * RY + r + 1;
            sa[b][lp][tx + 1 + hxd] = cur.hx[r];       // halo column on edge lanes (own cell otherwise,
            sa[b][lp][tx + 1] = tc[r];                 // overwritten by this lane's next write)
        }
        if (hyd < 0) sa[b][0][tx + 1] = cur.hy;
        if (hyd > 0) sa[b][ROWS + 1][tx + 1] = cur.hy;
        __syncthreads();
        // substep 1 at plane zs on the region
        const bool upd1 = !(zs == 0 || zs == a.Nz - 1);
        const BhCoef<T> c1 = coef(zs >= a.iz_tissue ? 1 : 0);      // after the barrier above
        T t1[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int lp = ty * RY + r + 1;
            const T u = bh_update_c(c1, tc[r], sa[b][lp][tx], sa[b][lp][tx + 2], sa[b][lp - 1][tx + 1],
                                    sa[b][lp + 1][tx + 1], cur.zp[r], tm[r], cur.q[r]);
            t1[r] = (upd1 && itr[r]) ? u : cur.bnd[r];
            sb[b][lp][tx + 1] = t1[r];
        }
        // substep 2 at plane zs - 1 on the inner cells (substep-1 plane zs-1
        // was written to sb[b^1] before this iteration's barrier)
        const int zo2 = zs - 1;
        if (zo2 >= z0) {
            const BhCoef<T> c2 = coef(zo2 >= a.iz_tissue ? 1 : 0);
            const size_t zo = (size_t)zo2 * plane;
            const int pb = b ^ 1;
            const rsrc_t rdst = make_rsrc(a.dst + zo, pbytes);
            const rsrc_t rd = make_rsrc(a.TD + zo, pbytes);
            const rsrc_t rc = make_rsrc(KAHAN ? a.TDc + zo : a.TD + zo, KAHAN ? pbytes : 0u);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int lp = ty * RY + r + 1;
                const T T2 = bh_update_c(c2, t1c[r], sb[pb][lp][tx], sb[pb][lp][tx + 2], sb[pb][lp - 1][tx + 1],
                                         sb[pb][lp + 1][tx + 1], t1[r], t1m[r], qp[r]);
                T td = hd[r], tdc = hdc[r];
                bh_accumulate(td, tdc, bh_dose_c(dc, T2), kahan);
                bstore<T>(T2, rdst, o_out[r]);
                if constexpr (KAHAN) bstore<T>(tdc, rc, o_out[r]);
                bstore<T>(td, rd, o_out[r]);
            }
        }
        // dose of substep 1 at plane zs, held until substep 2 reaches it
        // (lanes and planes without an output cell hold a value never stored)
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            T td = cur.td[r], tdc = cur.tdc[r];
            bh_accumulate(td, tdc, bh_dose_c(dc, t1[r]), kahan);
            hd[r] = td; hdc[r] = tdc;
            t1m[r] = t1c[r]; t1c[r] = t1[r];
            qp[r] = cur.q[r];
            tm[r] = tc[r]; tc[r] = cur.zp[r];
        }
    };

    P2 cur, nxt;
    load(zb, cur);
    for (int zs = zb; zs <= ze; ++zs) {
        body(zs, cur, nxt);
        cur = nxt;
    }
}

// Three substeps in one launch: bh_march2 with one more level.  The BX x
// TY*RY region runs substep 1 on every region cell, substep 2 one plane
// behind and substep 3 two planes behind; the output is the inner
// (BX-4) x (TY*RY-4) cells, and neighbouring regions overlap by four
// columns / rows (the region starts one cell left of and below its output
// block, so the first region reaches the column and row -1, which no kept
// value reads).  z: the workgroup owns output planes [z0, z1) and runs
// iterations zs = max(z0-2, 0) .. z1+1: substep 1 at zs (when zs <= Nz-1),
// substep 2 at zs-1 (planes z0-1 .. z1 that exist), substep 3 at zs-2
// (output planes).  Substep-1 and substep-2 values live in registers (own
// column, 3 planes each) and in LDS planes (x/y neighbours); the dose of the
// three substeps is added in order before TD is stored once.  A cell that
// a substep does not update (grid boundary planes, rows and columns) keeps
// its value at every level, read from `dst`: three single substeps
// alternate between the fields, whose boundary shells bd_bioheat_run makes
// equal, and the result goes to `dst`, where the third of them would leave
// it.  The per-cell arithmetic is that of three single substeps, so the
// result is bit-identical to them.
template <typename T, int BX, int TY, int RY, bool KAHAN>
__global__ void __launch_bounds__(BX * TY) bh_march3(BhArgs<T> a) {
    constexpr int ROWS = TY * RY;
    static_assert(TY >= 2 && RY >= 1 && BX % 64 == 0, "two thread rows of whole wavefronts");
    static_assert(ROWS >= 6, "region too short for three substeps");
    constexpr uint32_t OOB = 0xFFFFFFFFu;
    constexpr bool kahan = KAHAN;
    __shared__ T sa[2][ROWS + 2][BX + 2];    // T^n plane zs with a one-cell halo
    __shared__ T sb[2][ROWS + 2][BX + 2];    // substep-1 plane (its ring is never read into a kept value)
    __shared__ T sc[2][ROWS + 2][BX + 2];    // substep-2 plane (likewise)
    const int lb = xcd_remap((int)blockIdx.x, a.nblocks);
    const int zc = lb / (a.ntx * a.nty), t2i = lb % (a.ntx * a.nty);   // tile-major
    const int bx = t2i % a.ntx, by = t2i / a.ntx;
    const int tx = (int)threadIdx.x % BX, ty = (int)threadIdx.x / BX;   // ty < TY
    const int N = a.Nxy;
    const size_t plane = (size_t)N * N;
    const uint32_t ES = sizeof(T);
    const uint32_t pbytes = (uint32_t)(plane * ES);
    const int gx = bx * (BX - 4) + tx - 1;
    const int gy0 = by * (ROWS - 4) - 1;
    const int z0 = 1 + zc * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.Nz - 1);
    const BhDoseC<T> dc{a.dose_off, a.T_off, a.dt_T};
    __shared__ T scoef[2][6];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const BhCoef<T> c = bh_coef(a, m);
            scoef[m][0] = c.keep; scoef[m][1] = c.kxy; scoef[m][2] = c.kz;
            scoef[m][3] = c.dq; scoef[m][4] = c.perf; scoef[m][5] = c.theta;
        }
    }
    auto coef = [&](int m) {
        return BhCoef<T>{scoef[m][0], scoef[m][1], scoef[m][2], scoef[m][3], scoef[m][4], scoef[m][5]};
    };

    bool itr[RY];
    uint32_t o_in[RY], o_edge[RY], o_out[RY], cq[RY], ohx[RY];
    const int hxd = tx == 0 ? -1 : (tx == BX - 1 ? 1 : 0);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int ly = ty * RY + r;
        const int gy = gy0 + ly;
        const bool ing = gx >= 0 && gx < N && gy >= 0 && gy < N;
        itr[r] = gx >= 1 && gx <= N - 2 && gy >= 1 && gy <= N - 2;
        const bool outc = itr[r] && tx >= 2 && tx <= BX - 3 && ly >= 2 && ly <= ROWS - 3;
        const uint32_t o = ing ? ((uint32_t)gy * (uint32_t)N + (uint32_t)gx) * ES : OOB;
        o_in[r] = o;
        o_edge[r] = ing && !itr[r] ? o : OOB;
        o_out[r] = outc ? o : OOB;
        const int ixa = gx - a.shift_x, iya = gy - a.shift_y;
        const bool q_in = itr[r] && ixa >= 0 && ixa < N && iya >= 0 && iya < N;   // T:76
        cq[r] = q_in ? ((uint32_t)iya * (uint32_t)N + (uint32_t)ixa) * ES : OOB;
        const int hx = gx + hxd;
        ohx[r] = (hxd != 0 && gy >= 0 && gy < N && hx >= 0 && hx < N)
                     ? ((uint32_t)gy * (uint32_t)N + (uint32_t)hx) * ES : OOB;
    }
    const int hyd = ty == 0 ? -1 : (ty == TY - 1 ? 1 : 0);
    const int hyrow = hyd < 0 ? gy0 - 1 : gy0 + ROWS;
    const uint32_t ohy = (hyd != 0 && gx >= 0 && gx < N && hyrow >= 0 && hyrow < N)
                             ? ((uint32_t)hyrow * (uint32_t)N + (uint32_t)gx) * ES : OOB;

    struct P3 { T zp[RY], q[RY], td[RY], tdc[RY], bnd[RY], hx[RY], hy; };
    // loads of substep-1 plane zs (0 <= zs <= Nz-1); zp is plane zs+1 (0 past the last plane)
    auto load = [&](int zs, P3& v) {
        const size_t zo = (size_t)zs * plane;
        const rsrc_t rs = make_rsrc(a.src + zo, pbytes);
        const rsrc_t rn = make_rsrc(a.src + zo + plane, zs + 1 < a.Nz ? pbytes : 0u);
        const rsrc_t rq = make_rsrc(a.Q + zo, pbytes);
        const rsrc_t rb = make_rsrc(a.dst + zo, pbytes);
        const rsrc_t rd = make_rsrc(a.TD + zo, pbytes);
        const bool bplane = zs == 0 || zs == a.Nz - 1;
        const bool outp = zs >= z0 && zs < z1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const uint32_t od = outp ? o_out[r] : OOB;
            const uint32_t ob = bplane ? o_in[r] : o_edge[r];
            v.zp[r] = bload<T>(rn, o_in[r]);
            v.q[r] = bload<T>(rq, cq[r]);
            v.bnd[r] = ob != OOB ? bload<T>(rb, ob) : T(0);
            v.td[r] = bload<T>(rd, od);
            v.hx[r] = ohx[r] != OOB ? bload<T>(rs, ohx[r]) : T(0);
        }
        if constexpr (KAHAN) {
            const rsrc_t rc = make_rsrc(a.TDc + zo, pbytes);
#pragma unroll
            for (int r = 0; r < RY; ++r) v.tdc[r] = bload<T>(rc, outp ? o_out[r] : OOB);
        } else {
#pragma unroll
            for (int r = 0; r < RY; ++r) v.tdc[r] = T(0);
        }
        v.hy = ohy != OOB ? bload<T>(rs, ohy) : T(0);
    };

    const int zb = max(z0 - 2, 0), ze = z1 + 1;     // iterations
    const int last1 = min(ze, a.Nz - 1);             // last substep-1 plane
    const int lo2 = max(z0 - 1, 0), hi2 = min(z1, a.Nz - 1);
    // own column: T^n at zs-1, zs; substep 1 at zs-2, zs-1; substep 2 at zs-3,
    // zs-2; Q at zs-1, zs-2; boundary values of plane zs-1; dose after
    // substep 1 (plane zs-1) and after substep 2 (plane zs-2)
    T tm[RY], tc[RY], s1m[RY], s1c[RY], s2m[RY], s2c[RY], qp[RY], qpp[RY], bndp[RY];
    T h1[RY], h1c[RY], h2[RY], h2c[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const rsrc_t rs = make_rsrc(a.src + (size_t)zb * plane, pbytes);
        const rsrc_t rm = make_rsrc(a.src + (size_t)(zb > 0 ? zb - 1 : 0) * plane, zb >= 1 ? pbytes : 0u);
        tm[r] = bload<T>(rm, o_in[r]);
        tc[r] = bload<T>(rs, o_in[r]);
        s1m[r] = s1c[r] = s2m[r] = s2c[r] = qp[r] = qpp[r] = bndp[r] = T(0);
        h1[r] = h1c[r] = h2[r] = h2c[r] = T(0);
    }

    P3 cur, nxt;
    load(zb, cur);
    for (int zs = zb; zs <= ze; ++zs) {
        if (zs + 1 <= last1) load(zs + 1, nxt);
        const int b = (zs - zb) & 1;
        const bool do1 = zs <= a.Nz - 1;
        if (do1) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int lp = ty * RY + r + 1;
                sa[b][lp][tx + 1 + hxd] = cur.hx[r];   // halo column on edge lanes (own cell otherwise,
                sa[b][lp][tx + 1] = tc[r];             // overwritten by this lane's next write)
            }
            if (hyd < 0) sa[b][0][tx + 1] = cur.hy;
            if (hyd > 0) sa[b][ROWS + 1][tx + 1] = cur.hy;
        }
        __syncthreads();
        // substep 1 at plane zs on the region
        T s1[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) s1[r] = T(0);
        if (do1) {
            const bool upd = !(zs == 0 || zs == a.Nz - 1);
            const BhCoef<T> c1 = coef(zs >= a.iz_tissue ? 1 : 0);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int lp = ty * RY + r + 1;
                const T u = bh_update_c(c1, tc[r], sa[b][lp][tx], sa[b][lp][tx + 2], sa[b][lp - 1][tx + 1],
                                        sa[b][lp + 1][tx + 1], cur.zp[r], tm[r], cur.q[r]);
                s1[r] = (upd && itr[r]) ? u : cur.bnd[r];
                sb[b][lp][tx + 1] = s1[r];
            }
        }
        // substep 2 at plane zs-1 on the region (its neighbours: sb[b^1], written
        // by the previous iteration before this iteration's barrier)
        const int p2 = zs - 1;
        T s2[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) s2[r] = T(0);
        if (p2 >= lo2 && p2 <= hi2) {
            const bool upd = !(p2 == 0 || p2 == a.Nz - 1);
            const BhCoef<T> c2 = coef(p2 >= a.iz_tissue ? 1 : 0);
            const int pb = b ^ 1;
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int lp = ty * RY + r + 1;
                T v = bndp[r];
                if (upd) {
                    const T u = bh_update_c(c2, s1c[r], sb[pb][lp][tx], sb[pb][lp][tx + 2], sb[pb][lp - 1][tx + 1],
                                            sb[pb][lp + 1][tx + 1], s1[r], s1m[r], qp[r]);
                    if (itr[r]) v = u;
                }
                s2[r] = v;
                sc[b][lp][tx + 1] = v;
            }
        }
        // substep 3 at plane zs-2 on the output cells (neighbours: sc[b^1])
        const int p3 = zs - 2;
        if (p3 >= z0 && p3 < z1) {
            const BhCoef<T> c3 = coef(p3 >= a.iz_tissue ? 1 : 0);
            const size_t zo = (size_t)p3 * plane;
            const int pb = b ^ 1;
            const rsrc_t rdst = make_rsrc(a.dst + zo, pbytes);
            const rsrc_t rd = make_rsrc(a.TD + zo, pbytes);
            const rsrc_t rc = make_rsrc(KAHAN ? a.TDc + zo : a.TD + zo, KAHAN ? pbytes : 0u);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int lp = ty * RY + r + 1;
                const T T3 = bh_update_c(c3, s2c[r], sc[pb][lp][tx], sc[pb][lp][tx + 2], sc[pb][lp - 1][tx + 1],
                                         sc[pb][lp + 1][tx + 1], s2[r], s2m[r], qpp[r]);
                T td = h2[r], tdc = h2c[r];
                bh_accumulate(td, tdc, bh_dose_c(dc, T3), kahan);
                bstore<T>(T3, rdst, o_out[r]);
                if constexpr (KAHAN) bstore<T>(tdc, rc, o_out[r]);
                bstore<T>(td, rd, o_out[r]);
            }
        }
        // doses held until the next substep reaches the plane (lanes and
        // planes without an output cell hold values never stored)
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            T d2 = h1[r], d2c = h1c[r];
            bh_accumulate(d2, d2c, bh_dose_c(dc, s2[r]), kahan);
            h2[r] = d2; h2c[r] = d2c;
            T d1 = cur.td[r], d1c = cur.tdc[r];
            bh_accumulate(d1, d1c, bh_dose_c(dc, s1[r]), kahan);
            h1[r] = d1; h1c[r] = d1c;
            s2m[r] = s2c[r]; s2c[r] = s2[r];
            s1m[r] = s1c[r]; s1c[r] = s1[r];
            qpp[r] = qp[r]; qp[r] = cur.q[r];
            bndp[r] = cur.bnd[r];
            tm[r] = tc[r]; tc[r] = cur.zp[r];
        }
        if (zs + 1 <= last1) cur = nxt;
    }
}

static int first_tissue_plane_bh(const bd_bioheat_params* p, int Nz) {
    for (int iz = 0; iz < Nz; ++iz)
        if (!(iz * p->dz < p->water_tissue)) return iz;   // T:79
    return Nz;
}

template <typename T>
static int launch_bh(const bd_bioheat_params* p, const T* src, T* dst, const T* Q, T* TD, T* TDc, int Nxy,
                     int Nz, int sx, int sy, bool reference_order, hipStream_t st) {
    BhArgs<T> a;
    a.src = src; a.dst = dst; a.Q = Q; a.TD = TD; a.TDc = TDc;
    a.Nxy = Nxy; a.Nz = Nz;
    a.iz_tissue = first_tissue_plane_bh(p, Nz);
    a.shift_x = sx; a.shift_y = sy;
    const int nin = Nxy - 2;
    a.ntx = (nin + kBX - 1) / kBX;
    a.nty = (nin + kBY - 1) / kBY;
    const int nzin = Nz - 2;
    // target workgroup count (BD_BH_BLOCKS overrides it for the sweep)
    int target = 4096;
    if (const char* v = std::getenv("BD_BH_BLOCKS")) { const int x = std::atoi(v); if (x > 0) target = x; }
    int nzc = std::max(1, target / (a.ntx * a.nty));
    int zchunk = std::max(std::min(8, nzin), (nzin + nzc - 1) / nzc);
    a.nzc = (nzin + zchunk - 1) / zchunk;
    a.zchunk = zchunk;
    a.nblocks = a.ntx * a.nty * a.nzc;
    a.kxy[0] = (T)p->kdt_dxy_water; a.kxy[1] = (T)p->kdt_dxy;
    a.kz[0] = (T)p->kdt_dz_water; a.kz[1] = (T)p->kdt_dz;
    a.dq[0] = (T)p->dt_rho_ct_water; a.dq[1] = (T)p->dt_rho_ct;
    a.keep[0] = (T)(1.0 - 4.0 * p->kdt_dxy_water - 2.0 * p->kdt_dz_water);
    a.keep[1] = (T)(1.0 - 4.0 * p->kdt_dxy - 2.0 * p->kdt_dz);
    a.perf = (T)p->perf;
    a.theta_art = (T)(p->T_art - p->T_off);
    a.T_off = p->T_off;
    a.dt_T = p->dt_T;
    a.dose_off = (T)((p->T_off - 273.15) - 43.0);
    a.reference_order = reference_order ? 1 : 0;
    hipLaunchKernelGGL((bh_march<T>), dim3(a.nblocks), dim3(kBX * kBY), 0, st, a);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

// Region shapes of the pair kernel: BX columns (BX - 2 output columns), TY
// thread rows of RY rows each (TY * RY - 2 output rows).
struct Bh2Variant { int bx, ty, ry; };
constexpr Bh2Variant kBh2Variants[] = {{64, 4, 4}, {128, 4, 4}, {256, 4, 4}, {64, 4, 2}, {128, 4, 2}, {256, 4, 2}};
constexpr int kBh2NumVariants = (int)(sizeof(kBh2Variants) / sizeof(kBh2Variants[0]));
// FP64 keeps every value in registers with two rows per thread
constexpr bool bh2_f64_ok(int v) { return kBh2Variants[v].ry <= 2; }

template <typename T, int BX, int TY, int RY, bool KAHAN>
static int run_bh2(BhArgs<T> a, int variant, hipStream_t st) {
    constexpr int ROWS = TY * RY;
    const int nin = a.Nxy - 2;
    a.ntx = (nin + (BX - 2) - 1) / (BX - 2);
    a.nty = (nin + (ROWS - 2) - 1) / (ROWS - 2);
    const int nzin = a.Nz - 2;
    // z-chunk length: each workgroup runs zchunk + 2 substep-1 planes, and the
    // launch runs in ceil(workgroups / resident) rounds; take the length of
    // least rounds * (zchunk + 2), at least 8 planes (BD_BH_ZCHUNK overrides)
    constexpr int kMaxDev = 64;
    static int resident[kMaxDev][2][2][kBh2NumVariants] = {};
    static std::mutex mtx;
    int dev = 0, res = 0;
    BD_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDev) return BD_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lock(mtx);
        res = resident[dev][sizeof(T) == 8][KAHAN][variant];
    }
    if (res == 0) {
        int per_cu = 0;
        hipDeviceProp_t prop;
        BD_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        BD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bh_march2<T, BX, TY, RY, KAHAN>,
                                                                  BX * TY, 0));
        res = std::max(1, per_cu) * prop.multiProcessorCount;
        std::lock_guard<std::mutex> lock(mtx);
        resident[dev][sizeof(T) == 8][KAHAN][variant] = res;
    }
    const long long tiles = (long long)a.ntx * a.nty;
    int zchunk = nzin;
    long long best = -1;
    for (int zc = std::min(8, nzin); zc <= nzin; ++zc) {
        const int nzc = (nzin + zc - 1) / zc;
        if (zc > std::min(8, nzin) && (nzin + zc - 2) / (zc - 1) == nzc) continue;   // same split, longer chunk
        const long long rounds = (tiles * nzc + res - 1) / res;
        const long long cost = rounds * (zc + 2);
        if (best < 0 || cost < best) { best = cost; zchunk = zc; }
    }
    if (const char* v = std::getenv("BD_BH_ZCHUNK")) { const int x = std::atoi(v); if (x > 0) zchunk = std::min(x, nzin); }
    a.nzc = (nzin + zchunk - 1) / zchunk;
    a.zchunk = zchunk;
    a.nblocks = (int)(tiles * a.nzc);
    hipLaunchKernelGGL((bh_march2<T, BX, TY, RY, KAHAN>), dim3(a.nblocks), dim3(BX * TY), 0, st, a);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

template <typename T, int BX, int TY, int RY>
static int run_bh2_k(const BhArgs<T>& a, int variant, hipStream_t st) {
    return a.TDc ? run_bh2<T, BX, TY, RY, true>(a, variant, st) : run_bh2<T, BX, TY, RY, false>(a, variant, st);
}

// The pair kernel's region for a grid: rows of four (FP32) or two (FP64)
// per thread, and the width of 64, 128 or 256 columns that leaves the fewest
// idle lanes across the interior (the widest on a tie: fewer overlapping
// columns).  BD_BH_VARIANT overrides it, for the region sweep.
template <typename T>
static int bh2_variant(int Nxy) {
    const int base = sizeof(T) == 8 ? 3 : 0;
    const int nin = Nxy - 2;
    int v = base, best = -1;
    for (int k = 0; k < 3; ++k) {
        const int bx = kBh2Variants[base + k].bx;
        const int idle = (nin + bx - 3) / (bx - 2) * bx - nin;
        if (best < 0 || idle <= best) { best = idle; v = base + k; }
    }
    if (const char* e = std::getenv("BD_BH_VARIANT")) {
        const int x = std::atoi(e);
        if (x >= 0 && x < kBh2NumVariants && (sizeof(T) == 4 || bh2_f64_ok(x))) v = x;
    }
    return v;
}

// Region shapes of the triple kernel (FP32): BX columns (BX - 4 output
// columns), TY thread rows of two rows each (2 TY - 4 output rows).
struct Bh3Variant { int bx, ty; };
constexpr Bh3Variant kBh3Variants[] = {{64, 16}, {128, 8}};
constexpr int kBh3NumVariants = (int)(sizeof(kBh3Variants) / sizeof(kBh3Variants[0]));

template <int BX, int TY, bool KAHAN>
static int run_bh3(BhArgs<float> a, int variant, hipStream_t st) {
    constexpr int ROWS = TY * 2;
    const int nin = a.Nxy - 2;
    a.ntx = (nin + (BX - 4) - 1) / (BX - 4);
    a.nty = (nin + (ROWS - 4) - 1) / (ROWS - 4);
    const int nzin = a.Nz - 2;
    // z-chunk length: each workgroup runs zchunk + 3 iterations; the length
    // of least rounds * (zchunk + 3), at least 8 planes (BD_BH_ZCHUNK overrides)
    constexpr int kMaxDev = 64;
    static int resident[kMaxDev][2][kBh3NumVariants] = {};
    static std::mutex mtx;
    int dev = 0, res = 0;
    BD_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDev) return BD_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lock(mtx);
        res = resident[dev][KAHAN][variant];
    }
    if (res == 0) {
        int per_cu = 0;
        hipDeviceProp_t prop;
        BD_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        BD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bh_march3<float, BX, TY, 2, KAHAN>,
                                                                  BX * TY, 0));
        res = std::max(1, per_cu) * prop.multiProcessorCount;
        std::lock_guard<std::mutex> lock(mtx);
        resident[dev][KAHAN][variant] = res;
    }
    const long long tiles = (long long)a.ntx * a.nty;
    int zchunk = nzin;
    long long best = -1;
    for (int zc = std::min(8, nzin); zc <= nzin; ++zc) {
        const int nzc = (nzin + zc - 1) / zc;
        if (zc > std::min(8, nzin) && (nzin + zc - 2) / (zc - 1) == nzc) continue;   // same split, longer chunk
        const long long rounds = (tiles * nzc + res - 1) / res;
        const long long cost = rounds * (zc + 3);
        if (best < 0 || cost < best) { best = cost; zchunk = zc; }
    }
    if (const char* v = std::getenv("BD_BH_ZCHUNK")) { const int x = std::atoi(v); if (x > 0) zchunk = std::min(x, nzin); }
    a.nzc = (nzin + zchunk - 1) / zchunk;
    a.zchunk = zchunk;
    a.nblocks = (int)(tiles * a.nzc);
    hipLaunchKernelGGL((bh_march3<float, BX, TY, 2, KAHAN>), dim3(a.nblocks), dim3(BX * TY), 0, st, a);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

static BhArgs<float> bh_fused_args(const bd_bioheat_params* p, const float* src, float* dst, const float* Q,
                                   float* TD, float* TDc, int Nxy, int Nz, int sx, int sy);

// The triple kernel with region shape `variant` (index into kBh3Variants).
static int launch_bh3(const bd_bioheat_params* p, const float* src, float* dst, const float* Q, float* TD,
                      float* TDc, int Nxy, int Nz, int sx, int sy, int variant, hipStream_t st) {
    const BhArgs<float> a = bh_fused_args(p, src, dst, Q, TD, TDc, Nxy, Nz, sx, sy);
    if (variant == 0) return TDc ? run_bh3<64, 16, true>(a, 0, st) : run_bh3<64, 16, false>(a, 0, st);
    return TDc ? run_bh3<128, 8, true>(a, 1, st) : run_bh3<128, 8, false>(a, 1, st);
}

// Arguments of the fused kernels: bh_march2 and bh_march3 in the library's
// expression order.
template <typename T>
static BhArgs<T> bh_fused_args_t(const bd_bioheat_params* p, const T* src, T* dst, const T* Q, T* TD, T* TDc,
                                 int Nxy, int Nz, int sx, int sy) {
    BhArgs<T> a;
    a.src = src; a.dst = dst; a.Q = Q; a.TD = TD; a.TDc = TDc;
    a.Nxy = Nxy; a.Nz = Nz;
    a.iz_tissue = first_tissue_plane_bh(p, Nz);
    a.shift_x = sx; a.shift_y = sy;
    a.kxy[0] = (T)p->kdt_dxy_water; a.kxy[1] = (T)p->kdt_dxy;
    a.kz[0] = (T)p->kdt_dz_water; a.kz[1] = (T)p->kdt_dz;
    a.dq[0] = (T)p->dt_rho_ct_water; a.dq[1] = (T)p->dt_rho_ct;
    a.keep[0] = (T)(1.0 - 4.0 * p->kdt_dxy_water - 2.0 * p->kdt_dz_water);
    a.keep[1] = (T)(1.0 - 4.0 * p->kdt_dxy - 2.0 * p->kdt_dz);
    a.perf = (T)p->perf;
    a.theta_art = (T)(p->T_art - p->T_off);
    a.T_off = p->T_off;
    a.dt_T = p->dt_T;
    a.dose_off = (T)((p->T_off - 273.15) - 43.0);
    a.reference_order = 0;
    return a;
}

static BhArgs<float> bh_fused_args(const bd_bioheat_params* p, const float* src, float* dst, const float* Q,
                                   float* TD, float* TDc, int Nxy, int Nz, int sx, int sy) {
    return bh_fused_args_t<float>(p, src, dst, Q, TD, TDc, Nxy, Nz, sx, sy);
}

template <typename T>
static int launch_bh2(const bd_bioheat_params* p, const T* src, T* dst, const T* Q, T* TD, T* TDc, int Nxy,
                      int Nz, int sx, int sy, hipStream_t st) {
    const BhArgs<T> a = bh_fused_args_t<T>(p, src, dst, Q, TD, TDc, Nxy, Nz, sx, sy);
    const int v = bh2_variant<T>(Nxy);
    if constexpr (sizeof(T) == 4) {
        switch (v) {
            case 0: return run_bh2_k<T, 64, 4, 4>(a, v, st);
            case 1: return run_bh2_k<T, 128, 4, 4>(a, v, st);
            case 2: return run_bh2_k<T, 256, 4, 4>(a, v, st);
            case 3: return run_bh2_k<T, 64, 4, 2>(a, v, st);
            case 4: return run_bh2_k<T, 128, 4, 2>(a, v, st);
            default: return run_bh2_k<T, 256, 4, 2>(a, v, st);
        }
    } else {
        switch (v) {
            case 4: return run_bh2_k<T, 128, 4, 2>(a, v, st);
            case 5: return run_bh2_k<T, 256, 4, 2>(a, v, st);
            default: return run_bh2_k<T, 64, 4, 2>(a, v, st);
        }
    }
}

// dst's boundary shell (z planes 0 and Nz-1, rows 0 and N-1, columns 0 and
// N-1: the cells the update never changes) <- src's.  One thread per shell
// cell: two z planes, then two y rows and two x columns of each inner plane.
template <typename T>
__global__ void bh_shell_copy(const T* __restrict__ src, T* __restrict__ dst, int N, int Nz) {
    const long long zf = 2LL * N * N, per = 2LL * N + 2LL * (N - 2);
    const long long total = zf + per * (Nz - 2);
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    long long c;
    if (i < zf) {
        const long long k = i % ((long long)N * N);
        c = (i < (long long)N * N ? 0 : (long long)(Nz - 1) * N * N) + k;
    } else {
        const long long j = i - zf;
        const int iz = 1 + (int)(j / per), t = (int)(j % per);
        int ix, iy;
        if (t < 2 * N) { ix = t % N; iy = t < N ? 0 : N - 1; }
        else { const int u = t - 2 * N; iy = 1 + u % (N - 2); ix = u < N - 2 ? 0 : N - 1; }
        c = ((long long)iz * N + iy) * N + ix;
    }
    dst[c] = src[c];
}

static int check_bh(const bd_bioheat_params* p, int dtype, int Nxy, int Nz) {
    if (!p || (dtype != BD_F32 && dtype != BD_F64)) return BD_ERR_INVALID;
    if (Nxy < 3 || Nz < 3 || !(p->dz > 0) || !(p->dt_T > 0)) return BD_ERR_INVALID;
    if ((long long)Nxy * Nxy * 3 * 8 > 0xFFFFFFFFLL) return BD_ERR_INVALID;   // 32-bit in-plane offsets
    // explicit 7-point scheme: stable iff 1 - 4 kxy - 2 kz >= 0 in every region
    if (1.0 - 4.0 * p->kdt_dxy - 2.0 * p->kdt_dz - p->perf < 0.0) return BD_ERR_UNSTABLE;
    if (1.0 - 4.0 * p->kdt_dxy_water - 2.0 * p->kdt_dz_water < 0.0) return BD_ERR_UNSTABLE;
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_temperature_square(const bd_bioheat_params* prm, int dtype, void* T_n, void* T_p, const void* Q,
                                     void* TD, int Nxy, int Nz, int use_p, int shift_x, int shift_y, void* stream) {
    int rc = check_bh(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!T_n || !T_p || !Q || !TD || prm->T_off != 0.0) return BD_ERR_INVALID;
    void* dst = use_p ? T_n : T_p;
    const void* src = use_p ? T_p : T_n;
    hipStream_t st = as_stream(stream);
    if (dtype == BD_F32)
        return launch_bh<float>(prm, (const float*)src, (float*)dst, (const float*)Q, (float*)TD, nullptr, Nxy, Nz,
                                shift_x, shift_y, true, st);
    return launch_bh<double>(prm, (const double*)src, (double*)dst, (const double*)Q, (double*)TD, nullptr, Nxy,
                             Nz, shift_x, shift_y, true, st);
}

// n substeps ping-ponging between two fields, launched from here (no
// per-substep host round trip through the caller's language).
extern "C" int bd_bioheat_run(const bd_bioheat_params* prm, int dtype, void* T_a, void* T_b, const void* Q,
                              void* TD, void* TD_comp, int Nxy, int Nz, int shift_x, int shift_y, int n_steps,
                              int* which, void* stream) {
    int rc = check_bh(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!T_a || !T_b || !Q || !TD || T_a == T_b || !which || (*which != 0 && *which != 1) || n_steps < 0)
        return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    void* f[2] = {T_a, T_b};
    if (n_steps > 0) {
        // the other field takes the current field's boundary shell, so that
        // substeps paired or not see the same fixed boundary
        const long long shell = 2LL * Nxy * Nxy + (4LL * Nxy - 4) * (Nz - 2);
        const unsigned grid = (unsigned)((shell + 255) / 256);
        const int w = *which;
        if (dtype == BD_F32)
            hipLaunchKernelGGL(bh_shell_copy<float>, dim3(grid), dim3(256), 0, st, (const float*)f[w],
                               (float*)f[1 - w], Nxy, Nz);
        else
            hipLaunchKernelGGL(bh_shell_copy<double>, dim3(grid), dim3(256), 0, st, (const double*)f[w],
                               (double*)f[1 - w], Nxy, Nz);
        BD_LAUNCH_CHECK();
    }
    // three substeps in one launch (FP32), then pairs, then single substeps
    // (BD_BH_FUSE=2: pairs and single substeps; 0: single substeps only).
    // 256^3, 1,960 substeps (G cell-substeps/s, single / paired):
    // FP32 + Kahan 153 / 222, FP64 87 / 149.
    int fuse = dtype == BD_F32 ? 3 : 2;
    if (const char* v = std::getenv("BD_BH_FUSE")) {
        const int x = std::atoi(v);
        if (x == 0 || x == 2) fuse = std::min(fuse, x);
    }
    int k = 0;
    auto triple = [&](int v) {
        const int w = *which;
        const int r = launch_bh3(prm, (const float*)f[w], (float*)f[1 - w], (const float*)Q, (float*)TD,
                                 (float*)TD_comp, Nxy, Nz, shift_x, shift_y, v, st);
        if (!r) *which = 1 - w;
        return r;
    };
    auto pair = [&]() {
        const int w = *which;
        const int r = dtype == BD_F32
            ? launch_bh2<float>(prm, (const float*)f[w], (float*)f[1 - w], (const float*)Q, (float*)TD,
                                (float*)TD_comp, Nxy, Nz, shift_x, shift_y, st)
            : launch_bh2<double>(prm, (const double*)f[w], (double*)f[1 - w], (const double*)Q, (double*)TD,
                                 (double*)TD_comp, Nxy, Nz, shift_x, shift_y, st);
        if (!r) *which = 1 - w;
        return r;
    };
    // FP32: triples (any region shape) or pairs, whichever the grid runs
    // faster (their results are the same bits): measured once per device,
    // grid and dose mode on the first call with at least 11 substeps -- a
    // warm-up triple, then one launch of each timed with events -- and kept
    // for the process.  Measured (us per substep, pairs / 64 x 32 / 128 x 16
    // triples): 256^3 46 / 62 / 80, 426^2 x 300 201 / 162 / 193, 851^2 x 200
    // 482 / 431 / 391, 1024^2 x 100 390 / 320 / 357; a 64 x 16 region (512
    // threads) ran 60 / 177 / 430 / 324 and won nowhere.  BD_BH_VARIANT3
    // forces the triple shape.
    constexpr int kPairs = kBh3NumVariants;
    int mode = -1;                      // < kPairs: triples of that shape; kPairs: pairs
    if (fuse >= 3) {
        if (const char* e = std::getenv("BD_BH_VARIANT3")) {
            const int x = std::atoi(e);
            if (x >= 0 && x < kBh3NumVariants) mode = x;
        }
    }
    if (fuse >= 3 && mode < 0) {
        static std::mutex mtx;
        static std::map<std::tuple<int, int, int, int>, int> chosen;
        int dev = 0;
        BD_HIP_CHECK(hipGetDevice(&dev));
        const auto key = std::make_tuple(dev, Nxy, Nz, TD_comp ? 1 : 0);
        {
            std::lock_guard<std::mutex> lock(mtx);
            const auto it = chosen.find(key);
            if (it != chosen.end()) mode = it->second;
        }
        constexpr int kProbe = 3 + 3 * kBh3NumVariants + 2;    // warm-up triple, each shape, a pair
        if (mode < 0 && n_steps - k >= kProbe) {
            hipEvent_t ev[kPairs + 2] = {};
            int r = BD_OK;
            for (auto& e : ev) if (!r) r = hipEventCreate(&e) == hipSuccess ? BD_OK : BD_ERR_HIP;
            if (!r) r = triple(0);                               // warm-up
            for (int i = 0; i <= kPairs && !r; ++i) {
                r = hipEventRecord(ev[i], st) == hipSuccess ? BD_OK : BD_ERR_HIP;
                if (!r) r = i < kPairs ? triple(i) : pair();
            }
            if (!r) r = hipEventRecord(ev[kPairs + 1], st) == hipSuccess ? BD_OK : BD_ERR_HIP;
            if (!r) r = hipEventSynchronize(ev[kPairs + 1]) == hipSuccess ? BD_OK : BD_ERR_HIP;
            float per[kPairs + 1] = {};
            for (int i = 0; i <= kPairs && !r; ++i) {
                r = hipEventElapsedTime(&per[i], ev[i], ev[i + 1]) == hipSuccess ? BD_OK : BD_ERR_HIP;
                per[i] /= i < kPairs ? 3.f : 2.f;
            }
            for (auto& e : ev) if (e) (void)hipEventDestroy(e);
            if (r) return r;
            k += kProbe;
            mode = 0;
            for (int i = 1; i <= kPairs; ++i) if (per[i] < per[mode]) mode = i;
            std::lock_guard<std::mutex> lock(mtx);
            chosen[key] = mode;
        }
        if (mode < 0) mode = kPairs;
    }
    for (; fuse >= 3 && mode < kPairs && k + 3 <= n_steps; k += 3) {
        rc = triple(mode);
        if (rc) return rc;
    }
    for (; fuse >= 2 && k + 2 <= n_steps; k += 2) {
        rc = pair();
        if (rc) return rc;
    }
    for (; k < n_steps; ++k) {
        const int w = *which;
        if (dtype == BD_F32)
            rc = launch_bh<float>(prm, (const float*)f[w], (float*)f[1 - w], (const float*)Q, (float*)TD,
                                  (float*)TD_comp, Nxy, Nz, shift_x, shift_y, false, st);
        else
            rc = launch_bh<double>(prm, (const double*)f[w], (double*)f[1 - w], (const double*)Q, (double*)TD,
                                   (double*)TD_comp, Nxy, Nz, shift_x, shift_y, false, st);
        if (rc) return rc;
        *which = 1 - w;
    }
    return BD_OK;
}

extern "C" int bd_bioheat_step(const bd_bioheat_params* prm, int dtype, void* T_out, const void* T_in,
                               const void* Q, void* TD, void* TD_comp, int Nxy, int Nz, int shift_x, int shift_y,
                               void* stream) {
    int rc = check_bh(prm, dtype, Nxy, Nz);
    if (rc) return rc;
    if (!T_out || !T_in || !Q || !TD || T_out == T_in) return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    if (dtype == BD_F32)
        return launch_bh<float>(prm, (const float*)T_in, (float*)T_out, (const float*)Q, (float*)TD,
                                (float*)TD_comp, Nxy, Nz, shift_x, shift_y, false, st);
    return launch_bh<double>(prm, (const double*)T_in, (double*)T_out, (const double*)Q, (double*)TD,
                             (double*)TD_comp, Nxy, Nz, shift_x, shift_y, false, st);
}

// T:141-173, T:323-409 (host code): integer source shift per time step.
extern "C" int bd_spiral_reference(double speed, double scale, double speed_init, double t_acc, double dt,
                                   double dxy, double start, int spiral_lines, int* out_x, int* out_y, int cap) {
    if (!(speed > 0) || !(dt > 0) || !(dxy > 0) || !(t_acc > 0) || spiral_lines < 0 || cap < 0) return -1;
    if (cap > 0 && (!out_x || !out_y)) return -1;
    const int Nt_scale = 1 + (int)(scale / (speed * dt));                  // T:167
    const int Nt_acc = 1 + (int)(t_acc / dt);                              // T:170
    const double acceleration = (speed - speed_init) / t_acc;              // T:146
    std::vector<double> speed_acc(Nt_acc);
    for (int i = 0; i < Nt_acc; i++) speed_acc[i] = speed_init + i * dt * acceleration;  // T:172-173
    double x = start, y = start;
    int sgx = 1, sgy = -1, shx = 0, shy = 0;                               // T:334-336
    bool along_x = true;
    double xn = x, xp = x, yn = y, yp = y;
    double v = speed_init;
    int count = 0, grow = 1;
    for (int line = 0; line < spiral_lines; ++line) {
        const int steps = grow * Nt_scale + 2 * Nt_acc;
        xn = xp + sgx * dxy;
        yn = yp + sgy * dxy;
        for (int n = (line == 0 ? 1 : 0); n < steps; ++n) {
            // speed profile: ramp up, cruise, ramp down (T:357-362)
            if (n < Nt_acc) v = speed_acc[n];
            if (n >= Nt_acc && n < grow * Nt_scale + Nt_acc) v = speed;
            if (n >= grow * Nt_scale + Nt_acc) v = speed_acc[steps - n - 1];
            // move and snap to the nearer node (T:364-387)
            double& s = along_x ? x : y;
            double& prev = along_x ? xp : yp;
            double& next = along_x ? xn : yn;
            int& sh = along_x ? shx : shy;
            const int sg = along_x ? sgx : sgy;
            s = s + sg * v * dt;
            if (std::fabs(next - s) / dxy < std::fabs(s - prev) / dxy) {
                sh = sh + sg;
                prev = next;
                next = next + sg * dxy;
            }
            if (count < cap) { out_x[count] = shx; out_y[count] = shy; }
            ++count;
        }
        along_x = !along_x;                                                // T:394-406
        if (along_x) { ++grow; sgx = -sgx; xn = xp; }
        else { sgy = -sgy; yn = yp; }
    }
    return count;
}

extern "C" int bd_spiral_dwell(int n_side, double pitch, double dxy, int* sx, int* sy) {
    if (n_side < 1 || n_side % 2 == 0 || !(pitch > 0) || !(dxy > 0) || !sx || !sy) return BD_ERR_INVALID;
    // legs 1,1,2,2,...,(n-1),(n-1),(n-1): +x, +y, -x, -y, ... from the centre
    int gx = 0, gy = 0, k = 0;
    auto put = [&]() {
        sx[k] = (int)std::lround(gx * pitch / dxy);
        sy[k] = (int)std::lround(gy * pitch / dxy);
        ++k;
    };
    put();
    int len = 1, dir = 0;
    const int total = n_side * n_side;
    while (k < total) {
        for (int rep = 0; rep < 2 && k < total; ++rep) {
            for (int s = 0; s < len && k < total; ++s) {
                if (dir == 0) ++gx; else if (dir == 1) ++gy; else if (dir == 2) --gx; else --gy;
                put();
            }
            dir = (dir + 1) % 4;
        }
        ++len;
    }
    return BD_OK;
}

// Field slices in the reference's text format (T:412-447): a header line,
// then one line per row, each value followed by ';', numbers as the default
// stream format (printf %g).  XZ: rows iz in [r0, r1), columns ix in [c0, c1)
// at iy = fixed ("Z=rows;X=columns"); XY: rows iy, columns ix at iz = fixed
// ("Y=rows;X=columns").  `offset` is added to every value (fields stored as
// T - T_off).
extern "C" int bd_write_slice(const char* path, int dtype, const void* field, int Nxy, int Nz, int plane, int fixed,
                              int r0, int r1, int c0, int c1, double offset, void* stream) {
    if (!path || !field || (dtype != BD_F32 && dtype != BD_F64) || Nxy <= 0 || Nz <= 0) return BD_ERR_INVALID;
    if (plane != BD_SLICE_XZ && plane != BD_SLICE_XY) return BD_ERR_INVALID;
    const int rmax = plane == BD_SLICE_XZ ? Nz : Nxy, fmax = plane == BD_SLICE_XZ ? Nxy : Nz;
    if (r0 < 0 || r1 > rmax || r0 >= r1 || c0 < 0 || c1 > Nxy || c0 >= c1 || fixed < 0 || fixed >= fmax)
        return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    const size_t es = dtype == BD_F32 ? 4 : 8, nc = (size_t)(c1 - c0);
    std::vector<unsigned char> buf(nc * (size_t)(r1 - r0) * es);
    for (int r = r0; r < r1; ++r) {
        const size_t iz = plane == BD_SLICE_XZ ? (size_t)r : (size_t)fixed;
        const size_t iy = plane == BD_SLICE_XZ ? (size_t)fixed : (size_t)r;
        const size_t off = (iz * Nxy + iy) * Nxy + (size_t)c0;
        BD_HIP_CHECK(hipMemcpyAsync(buf.data() + (size_t)(r - r0) * nc * es, (const unsigned char*)field + off * es,
                                    nc * es, hipMemcpyDeviceToHost, st));
    }
    BD_HIP_CHECK(hipStreamSynchronize(st));
    FILE* f = std::fopen(path, "w");
    if (!f) return BD_ERR_IO;
    std::fprintf(f, plane == BD_SLICE_XZ ? "Z=rows;X=columns" : "Y=rows;X=columns");
    for (int r = 0; r < r1 - r0; ++r) {
        std::fprintf(f, "\n");
        for (size_t c = 0; c < nc; ++c) {
            const size_t k = (size_t)r * nc + c;
            const double v = dtype == BD_F32 ? (double)((const float*)buf.data())[k] : ((const double*)buf.data())[k];
            std::fprintf(f, "%g;", v + offset);
        }
    }
    return std::fclose(f) == 0 ? BD_OK : BD_ERR_IO;
}
