// bsn_slabs.hip -- a grid's state packed into (and unpacked from) the array
// layouts of the reference's files, one slab of k-planes at a time, so that
// snapshot and checkpoint I/O stream through small pinned buffers instead of a
// whole-array host copy.
//
//   layout BSN_LAYOUT_FIELDS: part q (0..4) is field q of the active cells,
//     element ((kk * ny + j) * nx + i) of a slab [k0, k0 + nk) -- one Julia
//     (nx, ny, nz) array per field, as write_snapshot stores /gas/rho ...
//     (src/io.jl:36-40);
//   layout BSN_LAYOUT_REF: one part, every cell with ghosts, element
//     (((kk * Ny + j) * Nx + i) * 5 + q) -- the Julia (5, Nx, Ny, Nz) array that
//     save_checkpoint_uniform stores as /grid/U (src/checkpoint.jl:276).
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"
#include "bsn_slabs.hpp"

namespace bsn {

#define NG BSN_NG

__device__ __forceinline__ int64_t slab_cell(const GridDev& G, int i, int j, int k) {
    return ((int64_t)k * G.Ny + j) * G.pitch + BSN_XOFF + i;
}

// A slab has fewer than 2^31 elements (the host sizes it), so the element
// index is 32-bit.
template <bool PACK>
__global__ void __launch_bounds__(256) k_slab(GridDev G, int layout, int part, int k0, int nk,
                                              double* __restrict__ buf) {
    if (layout == BSN_LAYOUT_FIELDS) {
        const int nx = G.nx, ny = G.ny;
        const int total = nk * ny * nx;
        double* __restrict__ f = G.U + (int64_t)part * G.fs;
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
            const int i = e % nx, r = e / nx;
            const int j = r % ny, kk = r / ny;
            const int64_t c = slab_cell(G, NG + i, NG + j, NG + k0 + kk);
            if (PACK) buf[e] = f[c];
            else f[c] = buf[e];
        }
    } else {
        const int Nx = G.nx + 2 * NG, Ny = G.Ny;
        const int total = nk * Ny * Nx * 5;
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
            const int q = e % 5, c5 = e / 5;
            const int i = c5 % Nx, r = c5 / Nx;
            const int j = r % Ny, kk = r / Ny;
            const int64_t c = (int64_t)q * G.fs + slab_cell(G, i, j, k0 + kk);
            if (PACK) buf[e] = G.U[c];
            else G.U[c] = buf[e];
        }
    }
}

static int slab_blocks(int64_t n, int cus) {
    int64_t b = (n + 255) / 256;
    const int64_t cap = (int64_t)cus * 16;
    if (b > cap) b = cap;
    return (int)(b < 1 ? 1 : b);
}

hipError_t launch_slab(hipStream_t s, int cus, const GridDev& G, int pack, int layout, int part,
                       int k0, int nk, double* buf) {
    const int64_t n = layout == BSN_LAYOUT_FIELDS ? (int64_t)nk * G.ny * G.nx
                                                  : (int64_t)nk * G.Ny * (G.nx + 2 * NG) * 5;
    if (n <= 0) return hipSuccess;
    if (pack)
        hipLaunchKernelGGL(k_slab<true>, dim3(slab_blocks(n, cus)), dim3(256), 0, s, G, layout, part,
                           k0, nk, buf);
    else
        hipLaunchKernelGGL(k_slab<false>, dim3(slab_blocks(n, cus)), dim3(256), 0, s, G, layout, part,
                           k0, nk, buf);
    return hipGetLastError();
}

}  // namespace bsn
