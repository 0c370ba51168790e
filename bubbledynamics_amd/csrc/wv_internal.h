// Internal interface between the Westervelt forward kernels and the adjoint driver.
#pragma once
#include "common.h"

namespace bd {

// Explicit level pointers of one forward step (device memory, plan dtype).
struct WvStepPtrs {
    void* pn; const void* p; const void* pm1; const void* pm2; const void* pm3; const void* pm4;
    void* r1; const void* r1m1; const void* r1m2;
    void* r2; const void* r2m1; const void* r2m2;
};

// Objective sphere of the adjoint workload (cell units).
struct WvObjective {
    int cx, cy, cz;
    double r2;
};

// One full forward step (interior + source plane + all faces) producing
// p^{step+1} at time t_next.  With obj != null the per-block partials written
// to epart[(step*nblocks + block)*2 + {0,1}] are sum p^2 inside / outside the
// sphere (interior cells only); otherwise the energy partials.
int wv_forward_step(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz, const WvStepPtrs& q, int step,
                    double t_next, const double* amp, const double* phase, double* epart, const WvObjective* obj,
                    hipStream_t st);
int wv_forward_nblocks(const bd_westervelt_params* prm, int dtype, int Nxy, int Nz);
int wv_check_params(const bd_westervelt_params* p, int dtype, int Nxy, int Nz);

}  // namespace bd
