// bsn_fft.hip -- the zero-padded FFT convolution of the isolated Poisson solve
// (solve_poisson_isolated, src/self_gravity.jl:35-76) as five pruned passes
// over power-of-two padded axes, FP64 complex, in LDS:
//
//   X  forward along x of the ny*nz rows that hold density (the other rows of
//      the padded array are zero), real input read straight from the grid
//      state with its density floor, half spectrum out:    A[k<nz][j<ny][m<=Nx/2]
//   Y  forward along y, zero-padded from ny to Ny:         B[k<nz][jj<Ny][m]
//   Z  forward along z (zero-padded from nz), times the real kernel spectrum,
//      inverse along z, only k < nz kept, in place in B
//   Y' inverse along y, only j < ny kept:                  A[k<nz][j<ny][m]
//   X' inverse along x from the Hermitian half spectrum, only i < nx kept,
//      scaled by dV / (Nx Ny Nz):                          phi[k][j][i]
//
// Only the octant the solve reads or writes ever reaches HBM.  Each block
// holds 2048 complex values in LDS: R = 2048 / N rows (x passes) or TW = 2048 /
// N columns of consecutive m (y, z passes), transformed by radix-4 stages
// (pairs of radix-2 stages): decimation in time on bit-reversed input for the
// forward transforms, decimation in frequency (bit-reversed output) for the
// inverse ones.
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"
#include "bsn_fft.hpp"

namespace bsn {

#pragma clang fp contract(fast)

#define NG BSN_NG
#define FFT_LCE 11           // log2 of the complex values per block in LDS
#define FFT_CE (1 << FFT_LCE)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ int brev(int e, int logN) {
    return (int)(__brev((unsigned)e) >> (32 - logN));
}

// Twiddles exp(-2 pi i p / N), p < N / 2, into LDS.
__device__ __forceinline__ void load_tw(double2* tw, const double2* __restrict__ g, int N) {
    for (int p = threadIdx.x; p < (N >> 1); p += blockDim.x) tw[p] = g[p];
}

// 2^lc transforms of length N = 2^logN at buf[e * 2^lc + c]; DIT takes
// bit-reversed input and leaves natural order, DIF the reverse; INV
// conjugates the twiddles (no 1/N).  Radix-2 stages are fused in pairs into
// radix-4 butterflies (one plain radix-2 stage when log2 N is odd), which
// halves the passes over LDS.  Every index is a shift or a mask (all extents
// are powers of two).  Ends with a barrier.
template <bool INV>
__device__ __forceinline__ double2 twid(const double2* tw, int p) {
    double2 w = tw[p];
    if (INV) w.y = -w.y;
    return w;
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <bool INV, bool DIT>
__device__ void fft_lds(double2* buf, const double2* tw, int logN, int lc) {
    const int cm = (1 << lc) - 1;
    for (int done = 0; done < logN;) {
        if (logN - done >= 2) {
            const int nq = 1 << (lc + logN - 2);                  // quads
            const int ls = DIT ? done : logN - 1 - done;
            const int lh = DIT ? ls : ls - 1;                     // h = 2^lh
            for (int b = threadIdx.x; b < nq; b += blockDim.x) {
                const int c = b & cm, qd = b >> lc;
                const int grp = qd >> lh, pos = qd & ((1 << lh) - 1);
                const int i0 = (((grp << (lh + 2)) + pos) << lc) + c, st = 1 << (lh + lc);
                const double2 x0 = buf[i0], x1 = buf[i0 + st], x2 = buf[i0 + 2 * st], x3 = buf[i0 + 3 * st];
                double2 z0, z1, z2, z3;
                // one twiddle per quad: the other two are its square and its
                // product with W^(N/4) = -i (+i for the inverse)
                if (DIT) {
                    const double2 wb = twid<INV>(tw, pos << (logN - 2 - ls));
                    const double2 wa = cmul(wb, wb);
                    const double2 wc = INV ? make_double2(-wb.y, wb.x) : make_double2(wb.y, -wb.x);
                    double2 t = cmul(wa, x1);
                    const double2 y0 = cadd(x0, t), y1 = csub(x0, t);
                    t = cmul(wa, x3);
                    const double2 y2 = cadd(x2, t), y3 = csub(x2, t);
                    t = cmul(wb, y2);
                    z0 = cadd(y0, t); z2 = csub(y0, t);
                    t = cmul(wc, y3);
                    z1 = cadd(y1, t); z3 = csub(y1, t);
                } else {
                    const double2 wa = twid<INV>(tw, pos << (logN - 1 - ls));
                    const double2 wb = INV ? make_double2(-wa.y, wa.x) : make_double2(wa.y, -wa.x);
                    const double2 wc = cmul(wa, wa);
                    const double2 y0 = cadd(x0, x2), y2 = cmul(wa, csub(x0, x2));
                    const double2 y1 = cadd(x1, x3), y3 = cmul(wb, csub(x1, x3));
                    z0 = cadd(y0, y1); z1 = cmul(wc, csub(y0, y1));
                    z2 = cadd(y2, y3); z3 = cmul(wc, csub(y2, y3));
                }
                buf[i0] = z0; buf[i0 + st] = z1; buf[i0 + 2 * st] = z2; buf[i0 + 3 * st] = z3;
            }
            done += 2;
        } else {
            const int nb = 1 << (lc + logN - 1);
            const int ls = DIT ? done : logN - 1 - done;          // half = 2^ls
            for (int b = threadIdx.x; b < nb; b += blockDim.x) {
                const int c = b & cm, bb = b >> lc;
                const int grp = bb >> ls, pos = bb & ((1 << ls) - 1);
                const int i0 = (((grp << (ls + 1)) + pos) << lc) + c, i1 = i0 + (1 << (ls + lc));
                const double2 w = twid<INV>(tw, pos << (logN - 1 - ls));
                const double2 a = buf[i0], v = buf[i1];
                if (DIT) {
                    const double2 t = cmul(w, v);
                    buf[i0] = cadd(a, t);
                    buf[i1] = csub(a, t);
                } else {
                    buf[i0] = cadd(a, v);
                    buf[i1] = cmul(w, csub(a, v));
                }
            }
            done += 1;
        }
        __syncthreads();
    }
}

// X forward.  rho from the grid state (field 0 of the active cells, floored
// when floors != 0) or, when rho != nullptr, from an (nx, ny, nz) array.
__global__ void __launch_bounds__(256) k_fft_x_fwd(GridDev G, const double* __restrict__ rho,
                                                   int floors, FftGeom F, double2* __restrict__ A,
                                                   const double2* __restrict__ twx) {
    extern __shared__ double2 lds[];              // FFT_CE values, then N / 2 twiddles
    double2* buf = lds;
    double2* tw = lds + FFT_CE;
    const int N = F.N[0], logN = F.logN[0], lr = FFT_LCE - logN, R = 1 << lr, H = (N >> 1) + 1;
    const int rows = F.n[1] * F.n[2];
    const int r0 = blockIdx.x * R;
    load_tw(tw, twx, N);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int r = idx >> logN, e = idx & (N - 1);
        const int row = r0 + r;
        double v = 0.0;
        if (e < F.n[0] && row < rows) {
            const int ly = F.logN[1] - 1;                 // ny = Ny / 2
            const int j = row & ((1 << ly) - 1), k = row >> ly;
            if (rho) {
                v = rho[(int64_t)row * F.n[0] + e];
            } else {
                v = G.U[((int64_t)(NG + k) * G.Ny + NG + j) * G.pitch + BSN_XOFF + NG + e];
                if (floors && v < G.rho_floor) v = G.rho_floor;   // src/gpu_kernels.jl:26-46
            }
        }
        buf[(brev(e, logN) << lr) + r] = make_double2(v, 0.0);
    }
    __syncthreads();
    fft_lds<false, true>(buf, tw, logN, lr);
    // m < N/2, then the Nyquist column m = N/2
    for (int idx = threadIdx.x; idx < (R << (logN - 1)); idx += blockDim.x) {
        const int r = idx >> (logN - 1), m = idx & ((N >> 1) - 1);
        if (r0 + r < rows) A[(int64_t)(r0 + r) * H + m] = buf[(m << lr) + r];
    }
    for (int r = threadIdx.x; r < R; r += blockDim.x)
        if (r0 + r < rows) A[(int64_t)(r0 + r) * H + (N >> 1)] = buf[((N >> 1) << lr) + r];
}

// Y forward: block (m-range, k).
__global__ void __launch_bounds__(256) k_fft_y_fwd(FftGeom F, const double2* __restrict__ A,
                                                   double2* __restrict__ B,
                                                   const double2* __restrict__ twy) {
    extern __shared__ double2 lds[];              // FFT_CE values, then N / 2 twiddles
    double2* buf = lds;
    double2* tw = lds + FFT_CE;
    const int N = F.N[1], logN = F.logN[1], lw = FFT_LCE - logN, TW = 1 << lw;
    const int H = (F.N[0] >> 1) + 1;
    const int m0 = blockIdx.x * TW, k = blockIdx.y, ny = F.n[1];
    load_tw(tw, twy, N);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        double2 v = make_double2(0.0, 0.0);
        if (e < ny && m < H) v = A[((int64_t)k * ny + e) * H + m];
        buf[(brev(e, logN) << lw) + c] = v;
    }
    __syncthreads();
    fft_lds<false, true>(buf, tw, logN, lw);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        if (m < H) B[((int64_t)k * N + e) * H + m] = buf[idx];
    }
}

// Z: forward, product with the kernel spectrum, inverse; block (m-range, jj).
__global__ void __launch_bounds__(256) k_fft_z(FftGeom F, double2* __restrict__ B,
                                               const double* __restrict__ Kc,
                                               const double2* __restrict__ twz) {
    extern __shared__ double2 lds[];              // FFT_CE values, then N / 2 twiddles
    double2* buf = lds;
    double2* tw = lds + FFT_CE;
    const int N = F.N[2], logN = F.logN[2], lw = FFT_LCE - logN, TW = 1 << lw;
    const int H = (F.N[0] >> 1) + 1;
    const int Ny = F.N[1], nz = F.n[2];
    const int m0 = blockIdx.x * TW, jj = blockIdx.y;
    load_tw(tw, twz, N);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        double2 v = make_double2(0.0, 0.0);
        if (e < nz && m < H) v = B[((int64_t)e * Ny + jj) * H + m];
        buf[(brev(e, logN) << lw) + c] = v;
    }
    __syncthreads();
    fft_lds<false, true>(buf, tw, logN, lw);
    // the kernel is even along every axis, so its spectrum is real and
    // symmetric: stored for kz <= Nz/2, ky <= Ny/2 only
    const int ky = min(jj, Ny - jj);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        const int kz = min(e, N - e);
        const double kv = m < H ? Kc[((int64_t)kz * ((Ny >> 1) + 1) + ky) * H + m] : 0.0;
        buf[idx] = make_double2(buf[idx].x * kv, buf[idx].y * kv);
    }
    __syncthreads();
    fft_lds<true, false>(buf, tw, logN, lw);
    for (int idx = threadIdx.x; idx < (nz << lw); idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        if (m < H) B[((int64_t)e * Ny + jj) * H + m] = buf[(brev(e, logN) << lw) + c];
    }
}

// Y inverse: block (m-range, k); only j < ny kept.
__global__ void __launch_bounds__(256) k_fft_y_inv(FftGeom F, const double2* __restrict__ B,
                                                   double2* __restrict__ A,
                                                   const double2* __restrict__ twy) {
    extern __shared__ double2 lds[];              // FFT_CE values, then N / 2 twiddles
    double2* buf = lds;
    double2* tw = lds + FFT_CE;
    const int N = F.N[1], logN = F.logN[1], lw = FFT_LCE - logN, TW = 1 << lw;
    const int H = (F.N[0] >> 1) + 1;
    const int m0 = blockIdx.x * TW, k = blockIdx.y, ny = F.n[1];
    load_tw(tw, twy, N);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        buf[idx] = m < H ? B[((int64_t)k * N + e) * H + m] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<true, false>(buf, tw, logN, lw);
    for (int idx = threadIdx.x; idx < (ny << lw); idx += blockDim.x) {
        const int e = idx >> lw, c = idx & (TW - 1), m = m0 + c;
        if (m < H) A[((int64_t)k * ny + e) * H + m] = buf[(brev(e, logN) << lw) + c];
    }
}

// X inverse: the full spectrum from its Hermitian half, only i < nx kept,
// scaled, into phi (nx, ny, nz).
__global__ void __launch_bounds__(256) k_fft_x_inv(FftGeom F, const double2* __restrict__ A,
                                                   double* __restrict__ phi, double scale,
                                                   const double2* __restrict__ twx) {
    extern __shared__ double2 lds[];              // FFT_CE values, then N / 2 twiddles
    double2* buf = lds;
    double2* tw = lds + FFT_CE;
    const int N = F.N[0], logN = F.logN[0], lr = FFT_LCE - logN, R = 1 << lr, H = (N >> 1) + 1;
    const int rows = F.n[1] * F.n[2], nx = F.n[0];
    const int r0 = blockIdx.x * R;
    load_tw(tw, twx, N);
    for (int idx = threadIdx.x; idx < FFT_CE; idx += blockDim.x) {
        const int r = idx >> logN, e = idx & (N - 1);
        double2 v = make_double2(0.0, 0.0);
        if (r0 + r < rows) {
            const double2* a = A + (int64_t)(r0 + r) * H;
            if (e < H) v = a[e];
            else { v = a[N - e]; v.y = -v.y; }
        }
        buf[(e << lr) + r] = v;
    }
    __syncthreads();
    fft_lds<true, false>(buf, tw, logN, lr);
    const int lx = F.logN[0] - 1;                     // nx = N / 2
    for (int idx = threadIdx.x; idx < (R << lx); idx += blockDim.x) {
        const int r = idx >> lx, i = idx & (nx - 1);
        if (r0 + r < rows) phi[(int64_t)(r0 + r) * nx + i] = buf[(brev(i, logN) << lr) + r].x * scale;
    }
}

// Real part of the kernel spectrum for kz <= Nz/2, ky <= Ny/2.
__global__ void k_fft_pack_kernel(FftGeom F, const double2* __restrict__ kspec, double* __restrict__ Kc) {
    const int H = (F.N[0] >> 1) + 1, hy = (F.N[1] >> 1) + 1, hz = (F.N[2] >> 1) + 1;
    const int64_t total = (int64_t)hz * hy * H;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(e % H);
        const int64_t r = e / H;
        const int ky = (int)(r % hy), kz = (int)(r / hy);
        Kc[e] = kspec[((int64_t)kz * F.N[1] + ky) * H + m].x;
    }
}

#pragma clang fp contract(off)

static int log2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

bool fft_geom(int nx, int ny, int nz, FftGeom& F) {
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        const int N = 2 * n[a];
        if (n[a] < 1 || (N & (N - 1)) != 0 || N > FFT_CE || N < 2) return false;
        F.n[a] = n[a];
        F.N[a] = N;
        F.logN[a] = log2i(N);
    }
    return true;
}

hipError_t fft_pack_kernel(hipStream_t s, const FftGeom& F, const double2* kspec, double* Kc) {
    hipLaunchKernelGGL(k_fft_pack_kernel, dim3(1024), dim3(256), 0, s, F, kspec, Kc);
    return hipGetLastError();
}

hipError_t fft_convolve(hipStream_t s, const FftGeom& F, const GridDev& G, const double* rho,
                        int floors, double2* A, double2* B, const double* Kc,
                        const double2* const tw[3], double* phi, double scale) {
    const int H = (F.N[0] >> 1) + 1;
    const int rows = F.n[1] * F.n[2];
    const int Rx = FFT_CE / F.N[0];
    const int TWy = FFT_CE / F.N[1], TWz = FFT_CE / F.N[2];
    auto lds = [](int N) { return (size_t)(FFT_CE + N / 2) * sizeof(double2); };
    hipLaunchKernelGGL(k_fft_x_fwd, dim3((rows + Rx - 1) / Rx), dim3(256), lds(F.N[0]), s, G, rho,
                       floors, F, A, tw[0]);
    hipLaunchKernelGGL(k_fft_y_fwd, dim3((H + TWy - 1) / TWy, F.n[2]), dim3(256), lds(F.N[1]), s, F,
                       A, B, tw[1]);
    hipLaunchKernelGGL(k_fft_z, dim3((H + TWz - 1) / TWz, F.N[1]), dim3(256), lds(F.N[2]), s, F, B,
                       Kc, tw[2]);
    hipLaunchKernelGGL(k_fft_y_inv, dim3((H + TWy - 1) / TWy, F.n[2]), dim3(256), lds(F.N[1]), s, F,
                       B, A, tw[1]);
    hipLaunchKernelGGL(k_fft_x_inv, dim3((rows + Rx - 1) / Rx), dim3(256), lds(F.N[0]), s, F, A, phi,
                       scale, tw[0]);
    return hipGetLastError();
}

}  // namespace bsn
