// bsn_bodies.hip -- what the bodies (black holes) and the run's diagnostics
// read from the gas, as device reductions over a grid's active cells:
//
//   k_body_force    gas_force_on_bh for every body in one pass over the density
//                   (src/gravity_bh.jl:91-113);
//   k_body_accrete  the sums of accrete! over each body's sink sphere, one body
//                   per blockIdx.y over the box of cells that can lie inside its
//                   sink radius (src/sinks.jl:73-109);
//   k_body_diag     gas_energy_total, gas_kinetic_total, gas_momentum_total,
//                   gas_angular_momentum_total and bound_gas_mass in one pass
//                   over the five fields (src/diagnostics.jl:21-175);
//   k_body_final    the per-block partial sums, one block per sum segment.
//
// The per-cell terms follow the reference's operation order in FP64 without
// contraction, so a per-cell decision (e_tot <= 0 in bound_gas_mass, r >= r_sink
// and rho <= rho_sink_min in accrete!) is the reference's bit for bit; only the
// order of the final sums differs.  The reductions are deterministic: the block
// count depends only on the problem size, and every tree has a fixed shape.
//
// Blocks are 64 x 4: each wave walks whole (j, k) rows of active cells, its
// lanes over i, so loads are coalesced and a row costs one index division.
// Partial sums are stored one accumulator after another (part[q * rows + row]),
// so the final pass reads each accumulator's partials contiguously.
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"
#include "bsn_physics.hpp"
#include "bsn_bodies.hpp"

namespace bsn {

#define NG BSN_NG

// Row-major offset of active cell (i, j, k), 0-based, in a field.
__device__ __forceinline__ int64_t bq_row(const GridDev& G, int j, int k) {
    return ((int64_t)(NG + k) * G.Ny + NG + j) * G.pitch + BSN_XOFF + NG;
}

// Sum over the block of the first `nused` accumulators, written to
// part[q * stride + row].  Lane order and wave order are fixed.
template <int NACC>
__device__ __forceinline__ void block_sum(double (&acc)[NACC], int nused, double* __restrict__ part,
                                          int64_t row, int64_t stride) {
    __shared__ double red[BQ_WAVES][NACC];
    const int lane = threadIdx.x, wave = threadIdx.y;
#pragma unroll
    for (int q = 0; q < NACC; ++q) {
        if (q < nused) {
            double v = acc[q];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[wave][q] = v;
        }
    }
    __syncthreads();
    const int t = wave * 64 + lane;
    if (t < nused) part[t * stride + row] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// gas_force_on_bh (src/gravity_bh.jl:100-111): per cell and body
//   r2 = ddx^2 + ddy^2 + ddz^2 + eps^2, r3 = r2 sqrt(r2),
//   fac = M / r3 * rho * dV, F += fac * (ddx, ddy, ddz).
__global__ void __launch_bounds__(256) k_body_force(GridDev G, BodyQ Q, double* __restrict__ part) {
    double acc[BQ_NACC_FORCE];
#pragma unroll
    for (int q = 0; q < BQ_NACC_FORCE; ++q) acc[q] = 0.0;
    const int rows = G.ny * G.nz;
    for (int r = blockIdx.x * BQ_WAVES + threadIdx.y; r < rows; r += gridDim.x * BQ_WAVES) {
        const int j = r % G.ny, k = r / G.ny;
        const double* __restrict__ rho_row = G.U + bq_row(G, j, k);
        const double yc = Q.y0 + ((double)(j + 1) - 0.5) * Q.dy;
        const double zc = Q.z0 + ((double)(k + 1) - 0.5) * Q.dz;
        for (int i = threadIdx.x; i < G.nx; i += 64) {
            const double rho = rho_row[i];
            const double xc = Q.x0 + ((double)(i + 1) - 0.5) * Q.dx;
#pragma unroll
            for (int b = 0; b < BSN_MAX_BH; ++b) {
                if (b < Q.n) {
                    const double ddx = xc - Q.px[b], ddy = yc - Q.py[b], ddz = zc - Q.pz[b];
                    const double r2 = ((ddx * ddx + ddy * ddy) + ddz * ddz) + Q.eps[b] * Q.eps[b];
                    const double r3 = r2 * sqrt(r2);
                    const double fac = ((Q.mass[b] / r3) * rho) * Q.dV;
                    acc[3 * b + 0] += fac * ddx;
                    acc[3 * b + 1] += fac * ddy;
                    acc[3 * b + 2] += fac * ddz;
                }
            }
        }
    }
    block_sum<BQ_NACC_FORCE>(acc, 3 * Q.n, part, blockIdx.x, gridDim.x);
}

// accrete! sums (src/sinks.jl:93-103): cells with r < r_sink and
// rho > rho_sink_min add rho * rate and m * rate, rate = dV / t_sink * dt.
// The skips are written as the reference's (r >= rs, rho <= rho_sink_min), so
// NaN compares the same way.  Body b = blockIdx.y walks the rows of its box.
__global__ void __launch_bounds__(256) k_body_accrete(GridDev G, BodyQ Q, double* __restrict__ part) {
    const int b = blockIdx.y;
    double acc[BQ_NACC_ACCRETE] = {0.0, 0.0, 0.0, 0.0};
    const int ex = Q.ext[b][0], ey = Q.ext[b][1], ez = Q.ext[b][2];
    const int rows = ey * ez;
    const double rs = Q.rs[b], rate = Q.rate[b];
    for (int r = blockIdx.x * BQ_WAVES + threadIdx.y; r < rows; r += gridDim.x * BQ_WAVES) {
        const int j = Q.lo[b][1] + r % ey, k = Q.lo[b][2] + r / ey;
        const int64_t row = bq_row(G, j, k);
        const double yc = Q.y0 + ((double)(j + 1) - 0.5) * Q.dy;
        const double zc = Q.z0 + ((double)(k + 1) - 0.5) * Q.dz;
        for (int ii = threadIdx.x; ii < ex; ii += 64) {
            const int i = Q.lo[b][0] + ii;
            const double xc = Q.x0 + ((double)(i + 1) - 0.5) * Q.dx;
            const double ddx = xc - Q.px[b], ddy = yc - Q.py[b], ddz = zc - Q.pz[b];
            const double rr = sqrt((ddx * ddx + ddy * ddy) + ddz * ddz);
            if (rr >= rs) continue;
            const int64_t c = row + i;
            const double rho = G.U[c];
            if (rho <= Q.rho_sink_min) continue;
            acc[0] += rho * rate;
            acc[1] += G.U[G.fs + c] * rate;
            acc[2] += G.U[2 * G.fs + c] * rate;
            acc[3] += G.U[3 * G.fs + c] * rate;
        }
    }
    block_sum<BQ_NACC_ACCRETE>(acc, BQ_NACC_ACCRETE, part, (int64_t)b * gridDim.x + blockIdx.x,
                               (int64_t)gridDim.y * gridDim.x);
}

// Gas diagnostics (src/diagnostics.jl):
//   0 sum E                                       gas_energy_total      (:21-26)
//   1 sum 0.5 (mx^2 + my^2 + mz^2) / max(rho, 1e-30)   gas_kinetic_total  (:33-44)
//   2-4 sum m                                     gas_momentum_total    (:51-59)
//   5-7 sum r x m about the origin                gas_angular_momentum_total (:67-94)
//   8 sum of max(rho, 1e-30) over cells with e_kin + e_int + Phi_BH <= 0
//                                                 bound_gas_mass        (:137-175)
// each multiplied by dV on the host after the sum, as the reference does.
__global__ void __launch_bounds__(256) k_body_diag(GridDev G, BodyQ Q, double* __restrict__ part) {
    double acc[BQ_NACC_DIAG];
#pragma unroll
    for (int q = 0; q < BQ_NACC_DIAG; ++q) acc[q] = 0.0;
    const double g1 = Q.gamma - 1.0;
    const int rows = G.ny * G.nz;
    for (int r = blockIdx.x * BQ_WAVES + threadIdx.y; r < rows; r += gridDim.x * BQ_WAVES) {
        const int j = r % G.ny, k = r / G.ny;
        const int64_t row = bq_row(G, j, k);
        const double yc = Q.y0 + ((double)(j + 1) - 0.5) * Q.dy;
        const double zc = Q.z0 + ((double)(k + 1) - 0.5) * Q.dz;
        for (int i = threadIdx.x; i < G.nx; i += 64) {
            const int64_t c = row + i;
            const double rho = G.U[c], mx = G.U[G.fs + c], my = G.U[2 * G.fs + c];
            const double mz = G.U[3 * G.fs + c], E = G.U[4 * G.fs + c];
            const double xc = Q.x0 + ((double)(i + 1) - 0.5) * Q.dx;
            const double rho_s = jl_max(rho, 1e-30);
            const double ke = (0.5 * ((mx * mx + my * my) + mz * mz)) / rho_s;
            acc[0] += E;
            acc[1] += ke;
            acc[2] += mx;
            acc[3] += my;
            acc[4] += mz;
            acc[5] += yc * mz - zc * my;
            acc[6] += zc * mx - xc * mz;
            acc[7] += xc * my - yc * mx;
            const double P = jl_max(g1 * (E - rho_s * ke), 0.0);
            const double e_int = P / (g1 * rho_s);
            double phi = 0.0;
#pragma unroll
            for (int b = 0; b < BSN_MAX_BH; ++b) {
                if (b < Q.n) {
                    const double ddx = xc - Q.px[b], ddy = yc - Q.py[b], ddz = zc - Q.pz[b];
                    phi = phi - Q.mass[b] / sqrt(((ddx * ddx + ddy * ddy) + ddz * ddz) +
                                                 Q.eps[b] * Q.eps[b]);
                }
            }
            const double e_tot = (ke + e_int) + phi;
            acc[8] += e_tot <= 0.0 ? rho_s : 0.0;
        }
    }
    block_sum<BQ_NACC_DIAG>(acc, BQ_NACC_DIAG, part, blockIdx.x, gridDim.x);
}

// out[s * nacc + q] = sum over b < nblocks of part[q * stride + s * nblocks + b]:
// block (q, s); thread t adds b = t, t + 256, ... in order, then a fixed tree.
__global__ void __launch_bounds__(256) k_body_final(const double* __restrict__ part, int nblocks,
                                                    int64_t stride, int nacc,
                                                    double* __restrict__ out) {
    __shared__ double red[4];
    const int q = blockIdx.x, s = blockIdx.y;
    const double* __restrict__ p = part + q * stride + (int64_t)s * nblocks;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) v += p[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[s * nacc + q] = ((red[0] + red[1]) + red[2]) + red[3];
}

static int body_blocks(int64_t rows) {
    int64_t b = (rows + BQ_WAVES - 1) / BQ_WAVES;
    if (b > BQ_MAX_BLOCKS) b = BQ_MAX_BLOCKS;
    return (int)(b < 1 ? 1 : b);
}

static hipError_t final_sum(hipStream_t s, const double* part, int nused, int nseg, int nblocks,
                            int nacc, double* out) {
    hipLaunchKernelGGL(k_body_final, dim3(nused, nseg), dim3(256), 0, s, part, nblocks,
                       (int64_t)nseg * nblocks, nacc, out);
    return hipGetLastError();
}

hipError_t launch_body_force(hipStream_t s, int, const GridDev& G, const BodyQ& Q, double* part,
                             double* out) {
    if (Q.n < 1) return hipSuccess;
    const int nb = body_blocks((int64_t)G.ny * G.nz);
    hipLaunchKernelGGL(k_body_force, dim3(nb), dim3(64, BQ_WAVES), 0, s, G, Q, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return final_sum(s, part, 3 * Q.n, 1, nb, BQ_NACC_FORCE, out);
}

hipError_t launch_body_accrete(hipStream_t s, int, const GridDev& G, const BodyQ& Q, double* part,
                               double* out) {
    if (Q.n < 1) return hipSuccess;
    int64_t rows = 0;
    for (int b = 0; b < Q.n; ++b) {
        const int64_t r = (int64_t)Q.ext[b][1] * Q.ext[b][2];
        if (r > rows) rows = r;
    }
    const int nb = body_blocks(rows);
    hipLaunchKernelGGL(k_body_accrete, dim3(nb, Q.n), dim3(64, BQ_WAVES), 0, s, G, Q, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return final_sum(s, part, BQ_NACC_ACCRETE, Q.n, nb, BQ_NACC_ACCRETE, out);
}

hipError_t launch_body_diag(hipStream_t s, int, const GridDev& G, const BodyQ& Q, double* part,
                            double* out) {
    const int nb = body_blocks((int64_t)G.ny * G.nz);
    hipLaunchKernelGGL(k_body_diag, dim3(nb), dim3(64, BQ_WAVES), 0, s, G, Q, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return final_sum(s, part, BQ_NACC_DIAG, 1, nb, BQ_NACC_DIAG, out);
}

}  // namespace bsn
