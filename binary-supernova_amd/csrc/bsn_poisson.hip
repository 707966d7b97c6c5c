// bsn_poisson.hip -- isolated-boundary Poisson solve for the gas self-gravity
// (solve_poisson_isolated, src/self_gravity.jl:35-76) on the device.
//
// Hockney-Eastwood zero padding: the active-cell density is placed in the
// first octant of a (2nx, 2ny, 2nz) array, convolved by FFT with the softened
// kernel K(r) = -1/sqrt(r^2 + eps^2), eps^2 = (dx dy dz)^(2/3), and the first
// octant of dV * IFFT(FFT(K) FFT(rho_pad)) is the potential.  The kernel
// spectrum is computed once per grid and cell size with real-to-complex hipFFT
// plans of the padded array (x fastest).  When every padded axis is a power of
// two the solves run the pruned transforms of bsn_fft.hip, which touch
// only the octant that holds density; otherwise they use the hipFFT plans.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <cmath>
#include <vector>
#include "bsn_internal.hpp"
#include "bsn_physics.hpp"
#include "bsn_poisson.hpp"

namespace bsn {

#define NG BSN_NG

// K on the padded grid, folded so that index 0 is the origin along each axis.
__global__ void k_green(double* __restrict__ pad, int nx, int ny, int nz, double dx, double dy,
                        double dz, double eps2) {
    const int Nx = 2 * nx, Ny = 2 * ny;
    const int64_t total = (int64_t)Nx * Ny * (2 * nz);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Nx);
        const int64_t r = e / Nx;
        const int j = (int)(r % Ny), k = (int)(r / Ny);
        const double fi = (double)(i < nx ? i : i - Nx);
        const double fj = (double)(j < ny ? j : j - Ny);
        const double fk = (double)(k < nz ? k : k - 2 * nz);
        const double r2 = (fi * dx) * (fi * dx) + (fj * dy) * (fj * dy) + (fk * dz) * (fk * dz);
        pad[e] = -1.0 / sqrt(r2 + eps2);
    }
}

// Zero-padded density from a grid state (field 0 of U; with floors, active
// cells are floored first, as the RHS does before the source is added).
__global__ void k_pad_state(GridDev G, double* __restrict__ pad, int floors) {
    const int Nx = 2 * G.nx, Ny = 2 * G.ny;
    const int64_t total = (int64_t)Nx * Ny * (2 * G.nz);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Nx);
        const int64_t r = e / Nx;
        const int j = (int)(r % Ny), k = (int)(r / Ny);
        double v = 0.0;
        if (i < G.nx && j < G.ny && k < G.nz) {
            const int64_t c = ((int64_t)(NG + k) * G.Ny + NG + j) * G.pitch + BSN_XOFF + NG + i;
            v = G.U[c];
            // the density floor resets rho < rho_floor (src/gpu_kernels.jl:26-46)
            if (floors && v < G.rho_floor) v = G.rho_floor;
        }
        pad[e] = v;
    }
}

// Zero-padded density from an (nx, ny, nz) array (x fastest).
__global__ void k_pad_array(const double* __restrict__ rho, double* __restrict__ pad, int nx,
                            int ny, int nz) {
    const int Nx = 2 * nx, Ny = 2 * ny;
    const int64_t total = (int64_t)Nx * Ny * (2 * nz);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Nx);
        const int64_t r = e / Nx;
        const int j = (int)(r % Ny), k = (int)(r / Ny);
        pad[e] = (i < nx && j < ny && k < nz) ? rho[((int64_t)k * ny + j) * nx + i] : 0.0;
    }
}

// spec *= kspec * scale  (scale = dV / number of padded cells: the inverse
// transform is unnormalised)
__global__ void k_cmul(hipfftDoubleComplex* __restrict__ spec,
                       const hipfftDoubleComplex* __restrict__ kspec, int64_t n, double scale) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const hipfftDoubleComplex a = spec[e], b = kspec[e];
        hipfftDoubleComplex c;
        c.x = (a.x * b.x - a.y * b.y) * scale;
        c.y = (a.x * b.y + a.y * b.x) * scale;
        spec[e] = c;
    }
}

// First octant of the padded result into phi (nx, ny, nz), x fastest.
__global__ void k_extract(const double* __restrict__ pad, double* __restrict__ phi, int nx, int ny,
                          int nz) {
    const int64_t total = (int64_t)nx * ny * nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % nx);
        const int64_t r = e / nx;
        const int j = (int)(r % ny), k = (int)(r / ny);
        phi[e] = pad[((int64_t)k * (2 * ny) + j) * (2 * nx) + i];
    }
}

static int blocks_for(int64_t n, int cus) {
    int64_t b = (n + 255) / 256;
    const int64_t cap = (int64_t)cus * 8;
    if (b > cap) b = cap;
    return (int)(b < 1 ? 1 : b);
}

#define FFTCHK(x) do { if ((x) != HIPFFT_SUCCESS) return hipErrorUnknown; } while (0)
#define HIPRET(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

hipError_t poisson_create(PoissonCtx& P, int nx, int ny, int nz, double dx, double dy, double dz,
                          hipStream_t stream, int cus) {
    P.nx = nx; P.ny = ny; P.nz = nz;
    P.dx = dx; P.dy = dy; P.dz = dz;
    P.cus = cus;
    const int Nx = 2 * nx, Ny = 2 * ny, Nz = 2 * nz;
    const int64_t nreal = (int64_t)Nx * Ny * Nz;
    P.ncplx = (int64_t)Nz * Ny * (Nx / 2 + 1);
    HIPRET(hipMalloc((void**)&P.pad, nreal * sizeof(double)));
    HIPRET(hipMalloc((void**)&P.spec, P.ncplx * sizeof(hipfftDoubleComplex)));
    HIPRET(hipMalloc((void**)&P.kspec, P.ncplx * sizeof(hipfftDoubleComplex)));
    size_t wf = 0, wi = 0;
    FFTCHK(hipfftCreate(&P.fwd));
    FFTCHK(hipfftCreate(&P.inv));
    FFTCHK(hipfftSetAutoAllocation(P.fwd, 0));
    FFTCHK(hipfftSetAutoAllocation(P.inv, 0));
    // slowest dimension first: the padded array is [z][y][x]
    FFTCHK(hipfftMakePlan3d(P.fwd, Nz, Ny, Nx, HIPFFT_D2Z, &wf));
    FFTCHK(hipfftMakePlan3d(P.inv, Nz, Ny, Nx, HIPFFT_Z2D, &wi));
    P.work_bytes = wf > wi ? wf : wi;
    if (P.work_bytes) HIPRET(hipMalloc(&P.work, P.work_bytes));
    FFTCHK(hipfftSetWorkArea(P.fwd, P.work));
    FFTCHK(hipfftSetWorkArea(P.inv, P.work));
    FFTCHK(hipfftSetStream(P.fwd, stream));
    FFTCHK(hipfftSetStream(P.inv, stream));
    // kernel spectrum (eps^2 = (dx dy dz)^(2/3), src/self_gravity.jl:45)
    const double eps2 = std::pow(dx * dy * dz, 2.0 / 3.0);
    hipLaunchKernelGGL(k_green, dim3(blocks_for(nreal, cus)), dim3(256), 0, stream, P.pad, nx, ny,
                       nz, dx, dy, dz, eps2);
    HIPRET(hipGetLastError());
    FFTCHK(hipfftExecD2Z(P.fwd, P.pad, P.kspec));
    if (fft_geom(nx, ny, nz, P.geom)) {
        const FftGeom& F = P.geom;
        const int64_t H = F.N[0] / 2 + 1;
        HIPRET(hipMalloc((void**)&P.A, (size_t)nz * ny * H * sizeof(double2)));
        HIPRET(hipMalloc((void**)&P.B, (size_t)nz * F.N[1] * H * sizeof(double2)));
        HIPRET(hipMalloc((void**)&P.Kc, (size_t)(F.N[2] / 2 + 1) * (F.N[1] / 2 + 1) * H * sizeof(double)));
        for (int a = 0; a < 3; ++a) {
            const int N = F.N[a];
            std::vector<double2> t(N / 2);
            for (int p = 0; p < N / 2; ++p) {
                const double ang = 2.0 * M_PI * (double)p / (double)N;
                t[p] = make_double2(std::cos(ang), -std::sin(ang));
            }
            HIPRET(hipMalloc((void**)&P.tw[a], t.size() * sizeof(double2)));
            HIPRET(hipMemcpyAsync(P.tw[a], t.data(), t.size() * sizeof(double2),
                                  hipMemcpyHostToDevice, stream));
            HIPRET(hipStreamSynchronize(stream));
        }
        HIPRET(fft_pack_kernel(stream, F, (const double2*)P.kspec, P.Kc));
        HIPRET(hipStreamSynchronize(stream));
        // the solves need neither the plans nor the padded arrays any more
        (void)hipfftDestroy(P.fwd);
        (void)hipfftDestroy(P.inv);
        P.fwd = P.inv = 0;
        HIPRET(hipFree(P.pad));
        HIPRET(hipFree(P.spec));
        HIPRET(hipFree(P.kspec));
        HIPRET(hipFree(P.work));
        P.pad = nullptr; P.spec = P.kspec = nullptr; P.work = nullptr;
        P.pruned = true;
    }
    return hipStreamSynchronize(stream);
}

void poisson_destroy(PoissonCtx& P) {
    if (P.fwd) (void)hipfftDestroy(P.fwd);
    if (P.inv) (void)hipfftDestroy(P.inv);
    (void)hipFree(P.pad);
    (void)hipFree(P.spec);
    (void)hipFree(P.kspec);
    (void)hipFree(P.work);
    (void)hipFree(P.A);
    (void)hipFree(P.B);
    (void)hipFree(P.Kc);
    for (int a = 0; a < 3; ++a) (void)hipFree(P.tw[a]);
    P = PoissonCtx{};
}

// The convolution once the padded density is in P.pad.
static hipError_t poisson_convolve(PoissonCtx& P, double* phi, hipStream_t stream) {
    const int64_t nreal = (int64_t)8 * P.nx * P.ny * P.nz;
    FFTCHK(hipfftExecD2Z(P.fwd, P.pad, P.spec));
    const double scale = (P.dx * P.dy * P.dz) / (double)nreal;
    hipLaunchKernelGGL(k_cmul, dim3(blocks_for(P.ncplx, P.cus)), dim3(256), 0, stream, P.spec,
                       P.kspec, P.ncplx, scale);
    HIPRET(hipGetLastError());
    FFTCHK(hipfftExecZ2D(P.inv, P.spec, P.pad));
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    hipLaunchKernelGGL(k_extract, dim3(blocks_for(n, P.cus)), dim3(256), 0, stream, P.pad, phi,
                       P.nx, P.ny, P.nz);
    return hipGetLastError();
}

static double conv_scale(const PoissonCtx& P) {
    return (P.dx * P.dy * P.dz) / (double)((int64_t)8 * P.nx * P.ny * P.nz);
}

hipError_t poisson_solve_state(PoissonCtx& P, const GridDev& G, int floors, double* phi,
                               hipStream_t stream) {
    if (P.pruned)
        return fft_convolve(stream, P.geom, G, nullptr, floors, P.A, P.B, P.Kc, P.tw, phi,
                            conv_scale(P));
    const int64_t nreal = (int64_t)8 * P.nx * P.ny * P.nz;
    hipLaunchKernelGGL(k_pad_state, dim3(blocks_for(nreal, P.cus)), dim3(256), 0, stream, G, P.pad,
                       floors);
    HIPRET(hipGetLastError());
    return poisson_convolve(P, phi, stream);
}

hipError_t poisson_solve_array(PoissonCtx& P, const double* rho, double* phi, hipStream_t stream) {
    if (P.pruned) {
        GridDev none{};
        return fft_convolve(stream, P.geom, none, rho, 0, P.A, P.B, P.Kc, P.tw, phi, conv_scale(P));
    }
    const int64_t nreal = (int64_t)8 * P.nx * P.ny * P.nz;
    hipLaunchKernelGGL(k_pad_array, dim3(blocks_for(nreal, P.cus)), dim3(256), 0, stream, rho,
                       P.pad, P.nx, P.ny, P.nz);
    HIPRET(hipGetLastError());
    return poisson_convolve(P, phi, stream);
}

}  // namespace bsn
