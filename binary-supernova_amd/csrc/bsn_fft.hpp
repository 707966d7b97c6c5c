// bsn_fft.hpp -- pruned FFT convolution for the isolated Poisson solve
// on power-of-two padded axes (bsn_fft.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"

namespace bsn {

struct FftGeom {
    int n[3];       // active extents
    int N[3];       // padded extents 2n, powers of two <= 2048
    int logN[3];
};

// false when an axis cannot use the pruned path (the padded extent is not a
// power of two, or exceeds 2048)
bool fft_geom(int nx, int ny, int nz, FftGeom& F);
// the real part of a half spectrum [Nz][Ny][Nx/2+1] for kz <= Nz/2, ky <= Ny/2
hipError_t fft_pack_kernel(hipStream_t s, const FftGeom& F, const double2* kspec, double* Kc);
// phi = scale * (inverse of (Kc * forward of the zero-padded density)) on the
// first octant; the density comes from rho (an (nx, ny, nz) array) or, when
// rho is null, from field 0 of G (floored when floors != 0).  A holds
// nz*ny*(Nx/2+1) and B nz*Ny*(Nx/2+1) complex values; tw[a] holds
// exp(-2 pi i p / N[a]) for p < N[a]/2.
hipError_t fft_convolve(hipStream_t s, const FftGeom& F, const GridDev& G, const double* rho,
                        int floors, double2* A, double2* B, const double* Kc,
                        const double2* const tw[3], double* phi, double scale);

}  // namespace bsn
