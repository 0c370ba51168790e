// Westervelt -> bioheat hand-off: the reference's Q3D binary file and the
// quarter -> full mirror expansion of the heat source.
//
// Reference: Westervelt_relax_fft_Pressure_Energy.cu (W:line) writes
// planes iz = 2..Nz-1 of the quarter-domain Q as float64, x fastest, then y,
// then z (W:778-790); Westervelt_Temperature_SquareSpiral.cu (T:line) reads
// (Nxy_half+1)^2 * Nz values and maps full-grid (ix, iy) to quarter cell
// (|ix - Nxy_half|, |iy - Nxy_half|) (T:271-302).
#include "common.h"
#include <cstdio>
#include <vector>

namespace bd {

// Full grid (Nxy x Nxy x Nz) from the quarter grid (Nq x Nq x Nz, Nq = Nxy/2 + 1).
template <typename S, typename D>
__global__ void mirror_expand_kernel(const S* __restrict__ q, D* __restrict__ f, int Nq, int Nxy, int Nz) {
    const size_t n = (size_t)Nxy * Nxy * Nz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int half = Nxy / 2;
    const int ix = (int)(i % Nxy), iy = (int)((i / Nxy) % Nxy), iz = (int)(i / ((size_t)Nxy * Nxy));
    const int sx = ix < half ? half - ix : ix - half;         // T:291-294
    const int sy = iy < half ? half - iy : iy - half;         // T:285-288
    f[i] = (D)q[(size_t)iz * Nq * Nq + (size_t)sy * Nq + sx];
}

template <typename S, typename D>
static int launch_expand(const S* q, D* f, int Nq, int Nxy, int Nz, hipStream_t st) {
    const size_t n = (size_t)Nxy * Nxy * Nz;
    hipLaunchKernelGGL((mirror_expand_kernel<S, D>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, q, f, Nq,
                       Nxy, Nz);
    BD_LAUNCH_CHECK();
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_mirror_expand(int dtype, const void* Q_quarter, void* Q_full, int Nxy_full, int Nz, void* stream) {
    if ((dtype != BD_F32 && dtype != BD_F64) || !Q_quarter || !Q_full || Nxy_full < 3 || Nz < 1 ||
        Q_quarter == Q_full)
        return BD_ERR_INVALID;
    const int Nq = Nxy_full / 2 + 1;
    hipStream_t st = as_stream(stream);
    if (dtype == BD_F32) return launch_expand<float, float>((const float*)Q_quarter, (float*)Q_full, Nq, Nxy_full, Nz, st);
    return launch_expand<double, double>((const double*)Q_quarter, (double*)Q_full, Nq, Nxy_full, Nz, st);
}

extern "C" int bd_q3d_write(const char* path, int dtype, const void* Q, int Nxy, int Nz, int iz0, void* stream) {
    if (!path || !Q || (dtype != BD_F32 && dtype != BD_F64) || Nxy < 1 || iz0 < 0 || iz0 >= Nz) return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    const size_t plane = (size_t)Nxy * Nxy, n = plane * (size_t)(Nz - iz0);
    std::vector<double> h(n);
    if (dtype == BD_F64) {
        BD_HIP_CHECK(hipMemcpyAsync(h.data(), (const double*)Q + plane * iz0, n * sizeof(double),
                                    hipMemcpyDeviceToHost, st));
    } else {
        std::vector<float> hf(n);
        BD_HIP_CHECK(hipMemcpyAsync(hf.data(), (const float*)Q + plane * iz0, n * sizeof(float),
                                    hipMemcpyDeviceToHost, st));
        BD_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t k = 0; k < n; ++k) h[k] = (double)hf[k];
    }
    BD_HIP_CHECK(hipStreamSynchronize(st));
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return BD_ERR_IO;
    const size_t w = std::fwrite(h.data(), sizeof(double), n, fp);
    const int closed = std::fclose(fp);
    return (w == n && closed == 0) ? BD_OK : BD_ERR_IO;
}

extern "C" int bd_q3d_read_expand(const char* path, int dtype, void* Q_full, int Nxy_full, int Nz, void* stream) {
    if (!path || !Q_full || (dtype != BD_F32 && dtype != BD_F64) || Nxy_full < 3 || Nz < 1) return BD_ERR_INVALID;
    hipStream_t st = as_stream(stream);
    const int Nq = Nxy_full / 2 + 1;
    const size_t n = (size_t)Nq * Nq * Nz;
    std::vector<double> h(n);
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return BD_ERR_IO;
    const size_t r = std::fread(h.data(), sizeof(double), n, fp);
    std::fclose(fp);
    if (r != n) return BD_ERR_IO;
    double* dq = nullptr;
    BD_HIP_CHECK(hipMalloc(&dq, n * sizeof(double)));
    int rc = BD_OK;
    if (hipMemcpyAsync(dq, h.data(), n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) rc = BD_ERR_HIP;
    if (rc == BD_OK) {
        rc = dtype == BD_F64 ? launch_expand<double, double>(dq, (double*)Q_full, Nq, Nxy_full, Nz, st)
                             : launch_expand<double, float>(dq, (float*)Q_full, Nq, Nxy_full, Nz, st);
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == BD_OK) rc = BD_ERR_HIP;
    (void)hipFree(dq);
    return rc;
}
