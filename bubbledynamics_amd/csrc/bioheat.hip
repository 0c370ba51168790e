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

// Loads of one plane for one thread, issued a plane ahead of their use.
template <typename T>
struct BhPlane { T zp, hx, hy, q, td, tdc; };

template <typename T>
__global__ void __launch_bounds__(kBX * kBY) bh_march(BhArgs<T> a) {
    __shared__ T s[2][kBY + 2][kBX + 2];
    const int lb = xcd_remap((int)blockIdx.x, a.nblocks);
    const int zc = lb % a.nzc, t2 = lb / a.nzc;
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
            const T Tc = zc0;
            const T l = s[b][ty + 1][tx], r = s[b][ty + 1][tx + 2];
            const T d = s[b][ty][tx + 1], u = s[b][ty + 2][tx + 1];
            T Tn;
            if (a.reference_order) {
                Tn = Tc * (T(1.0) - T(4.0) * a.kxy[reg] - T(2.0) * a.kz[reg]) + a.kxy[reg] * (l + r + u + d) +
                     a.kz[reg] * (cur.zp + zm) + a.dq[reg] * cur.q;
            } else {
                Tn = Tc * a.keep[reg] + a.kxy[reg] * ((l + r) + (u + d)) + a.kz[reg] * (cur.zp + zm) +
                     a.dq[reg] * cur.q;
            }
            if (reg && a.perf != T(0)) Tn = Tn - a.perf * (Tc - a.theta_art);
            bstore<T>(Tn, make_rsrc(a.dst + zo, pbytes), o);
            T add;
            if constexpr (std::is_same<T, float>::value) {
                const float x = Tn + a.dose_off;               // T_c - 43
                add = exp2f(x < 0.0f ? 2.0f * x : x) * (float)a.dt_T;
            } else {
                const double T_c = (Tn + a.T_off) - 273.15;   // T:88, T:96
                add = (T_c < 43.0) ? pow(4.0, T_c - 43.0) * a.dt_T : pow(2.0, T_c - 43.0) * a.dt_T;  // T:99-102
            }
            const rsrc_t rtd = make_rsrc(a.TD + zo, pbytes);
            if (!a.TDc) {
                bstore<T>(cur.td + add, rtd, o);
            } else {
                const T y = add - cur.tdc, tt = cur.td + y;
                bstore<T>((tt - cur.td) - y, make_rsrc(a.TDc + zo, pbytes), o);
                bstore<T>(tt, rtd, o);
            }
        }
        zm = zc0;
        zc0 = cur.zp;
        cur = nxt;
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
    for (int k = 0; k < n_steps; ++k) {
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
