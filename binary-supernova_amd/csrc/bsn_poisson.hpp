// bsn_poisson.hpp -- device Poisson solve with isolated boundaries (host API).
#pragma once
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include "bsn_internal.hpp"
#include "bsn_fft.hpp"

namespace bsn {

struct PoissonCtx {
    int nx = 0, ny = 0, nz = 0, cus = 1;
    double dx = 0, dy = 0, dz = 0;
    hipfftHandle fwd = 0, inv = 0;
    double* pad = nullptr;                  // (2nz, 2ny, 2nx) real, x fastest
    hipfftDoubleComplex* spec = nullptr;    // its half spectrum
    hipfftDoubleComplex* kspec = nullptr;   // spectrum of the softened kernel
    int64_t ncplx = 0;
    void* work = nullptr;
    size_t work_bytes = 0;
    // pruned transforms when every padded axis is a power of two <= 2048
    // (the hipFFT plans then only form the kernel spectrum at set-up)
    bool pruned = false;
    FftGeom geom{};
    double2* A = nullptr;                   // nz*ny*(Nx/2+1) complex
    double2* B = nullptr;                   // nz*Ny*(Nx/2+1) complex
    double* Kc = nullptr;                   // real kernel spectrum, kz <= Nz/2, ky <= Ny/2
    double2* tw[3] = {nullptr, nullptr, nullptr};
};

hipError_t poisson_create(PoissonCtx& P, int nx, int ny, int nz, double dx, double dy, double dz,
                          hipStream_t stream, int cus);
void poisson_destroy(PoissonCtx& P);
// phi (nx, ny, nz, x fastest) from the density of a grid state; with floors,
// the stage-1 density floor is applied on load
hipError_t poisson_solve_state(PoissonCtx& P, const GridDev& G, int floors, double* phi,
                               hipStream_t stream);
// phi from a density array of the same shape
hipError_t poisson_solve_array(PoissonCtx& P, const double* rho, double* phi, hipStream_t stream);

}  // namespace bsn
