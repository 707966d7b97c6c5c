// bsn_fft.hip -- the zero-padded FFT convolution of the isolated Poisson solve
// (solve_poisson_isolated, src/self_gravity.jl:35-76) as five pruned passes
// over power-of-two padded axes, FP64 complex:
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
// Only the octant the solve reads or writes ever reaches HBM.  A pass whose
// padded length N = P * Q has Q = P or 2P and 16 <= N <= 256 runs the register
// four-step form: each transform lives in the registers of P threads (Q values
// each), one LDS exchange between a Q-point and a P-point DFT.  Other lengths
// hold 2048 complex values per block in LDS: R = 2048 / N rows (x passes) or
// TW = 2048 / N columns of consecutive m (y, z passes), transformed by radix-4
// stages (pairs of radix-2 stages): decimation in time on bit-reversed input
// for the forward transforms, decimation in frequency (bit-reversed output)
// for the inverse ones.
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

// ---------------------------------------------------------------------------
// Register-resident transforms (N = P * Q with Q = P or 2P, 16 <= N <= 256).  A length-N transform is done by P threads, each
// holding Q values: thread p holds element p + P*q in register q.  Four steps:
// a length-Q DFT in registers, twiddles W_N^(p s), one transpose through LDS,
// length-P DFTs in registers.  On exit register r*P + k1 of thread p holds
// element (p + P*r) + Q*k1, which reordered (register r + (Q/P)*k1) is again
// "element p + P*q in register q": a forward transform feeds the inverse one
// without another transpose.

static constexpr double W32R[16] = {1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452,
                                   0.7071067811865476, 0.5555702330196023, 0.38268343236508984,
                                   0.19509032201612833, 0.0, -0.1950903220161282, -0.3826834323650897,
                                   -0.555570233019602, -0.7071067811865475, -0.8314696123025453,
                                   -0.9238795325112867, -0.9807852804032304};
static constexpr double W32I[16] = {0.0, -0.19509032201612825, -0.3826834323650898, -0.5555702330196022,
                                   -0.7071067811865475, -0.8314696123025452, -0.9238795325112867,
                                   -0.9807852804032304, -1.0, -0.9807852804032304, -0.9238795325112867,
                                   -0.8314696123025455, -0.7071067811865476, -0.5555702330196022,
                                   -0.3826834323650899, -0.1950903220161286};

constexpr int brev_c(int i, int L) {
    int r = 0;
    for (int b = 1; b < L; b <<= 1, i >>= 1) r = (r << 1) | (i & 1);
    return r;
}

// w * v for the twiddle W_32^k (conjugated for INV), k a compile-time index
template <int K, bool INV>
__device__ __forceinline__ double2 tw32(double2 v) {
    if constexpr (K == 0) {
        return v;
    } else if constexpr (K == 8) {
        return INV ? make_double2(-v.y, v.x) : make_double2(v.y, -v.x);     // +i / -i
    } else {
        const double wr = W32R[K], wi = INV ? -W32I[K] : W32I[K];
        return make_double2(v.x * wr - v.y * wi, v.x * wi + v.y * wr);
    }
}

template <int L, bool INV, int HALF, int POS>
__device__ __forceinline__ void bfly_reg(double2* a) {
#pragma unroll
    for (int g = 0; g < L; g += 2 * HALF) {
        const double2 t = tw32<POS * (32 / (2 * HALF)), INV>(a[g + POS + HALF]);
        const double2 u = a[g + POS];
        a[g + POS] = make_double2(u.x + t.x, u.y + t.y);
        a[g + POS + HALF] = make_double2(u.x - t.x, u.y - t.y);
    }
}

template <int L, bool INV, int HALF, int POS = 0>
__device__ __forceinline__ void stage_reg(double2* a) {
    bfly_reg<L, INV, HALF, POS>(a);
    if constexpr (POS + 1 < HALF) stage_reg<L, INV, HALF, POS + 1>(a);
}

template <int L, bool INV, int HALF = 1>
__device__ __forceinline__ void stages_reg(double2* a) {
    stage_reg<L, INV, HALF>(a);
    if constexpr (2 * HALF < L) stages_reg<L, INV, 2 * HALF>(a);
}

// In-place DFT of a[0..L-1] (natural order in and out), L <= 32, radix-2 with
// constant twiddles.
template <int L, bool INV>
__device__ __forceinline__ void dft_reg(double2* a) {
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const int j = brev_c(i, L);
        if (j > i) { const double2 t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    stages_reg<L, INV>(a);
}

// W_N^m for 0 <= m < N from the table of W_N^p, p < N/2
template <bool INV>
__device__ __forceinline__ double2 twN(const double2* tw, int m, int N) {
    double2 w = tw[m & ((N >> 1) - 1)];
    if (m >= (N >> 1)) { w.x = -w.x; w.y = -w.y; }
    if (INV) w.y = -w.y;
    return w;
}

// The four steps on x (thread p, transform c of ncol); lds holds N * ncol
// values, transform-interleaved (ROWS = false: value i of transform c at
// i * ncol + c, for threads with c fastest) or transform by transform
// (ROWS = true: c * (N + Q) + i + i / P, for threads with p fastest; the pad
// of one slot per P keeps the strided reads of the second step off a single
// bank).
template <int P, int Q, bool INV, bool ROWS = false>
__device__ __forceinline__ void fft_reg(double2 (&x)[Q], int p, int c, int ncol, double* lds,
                                        const double2* tw) {
    constexpr int N = P * Q;
    auto at = [&](int i) -> double& {
        return ROWS ? lds[c * (N + Q) + i + i / P] : lds[i * ncol + c];
    };
    dft_reg<Q, INV>(x);
#pragma unroll
    for (int s = 0; s < Q; ++s) x[s] = cmul(x[s], twN<INV>(tw, p * s, N));
    // the exchange moves real parts, then imaginary parts, through one
    // FP64 buffer: half the LDS of a complex buffer, so twice the blocks
    // fit on a CU.  The real-part reads overwrite only .x, so .y still
    // holds this thread's own values when they are stored.
    __syncthreads();                       // lds free (a previous transform read it)
#pragma unroll
    for (int s = 0; s < Q; ++s) at(s * P + p) = x[s].x;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Q / P; ++r)
#pragma unroll
        for (int k = 0; k < P; ++k) x[r * P + k].x = at((p + P * r) * P + k);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < Q; ++s) at(s * P + p) = x[s].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Q / P; ++r) {
#pragma unroll
        for (int k = 0; k < P; ++k) x[r * P + k].y = at((p + P * r) * P + k);
        dft_reg<P, INV>(x + r * P);
    }
}

// Element index held by register j after fft_reg, and the reordering that
// makes the output the next transform's input.
template <int P, int Q>
__device__ __forceinline__ int reg_elem(int p, int j) {
    const int r = j / P, k1 = j % P;
    return (p + P * r) + Q * k1;
}
template <int P, int Q>
__device__ __forceinline__ void reg_to_input(double2 (&x)[Q]) {
    double2 t[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) t[(j / P) + (Q / P) * (j % P)] = x[j];
#pragma unroll
    for (int j = 0; j < Q; ++j) x[j] = t[j];
}

// Y forward, register form: 256 threads = (256/P columns) x P; the columns of
// a block are consecutive in the flat (k, m) order.
template <int P, int Q>
__global__ void __launch_bounds__(256) k_fftr_y_fwd(FftGeom F, const double2* __restrict__ A,
                                                    double2* __restrict__ B,
                                                    const double2* __restrict__ twy) {
    constexpr int N = P * Q, TW = 256 / P;
    extern __shared__ double2 lds[];                  // N * TW FP64 values, then N / 2 twiddles
    double2* tw = lds + N * TW / 2;
    const int H = (F.N[0] >> 1) + 1, ny = F.n[1];
    const int c = threadIdx.x % TW, p = threadIdx.x / TW;
    const int g = blockIdx.x * TW + c;                // flat column over (k, m)
    const bool ok = g < H * F.n[2];
    const int k = ok ? g / H : 0, m = ok ? g - k * H : 0;
    load_tw(tw, twy, N);
    double2 x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = p + P * q;
        x[q] = (q < Q / 2 && e < ny && ok) ? A[((int64_t)k * ny + e) * H + m] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_reg<P, Q, false>(x, p, c, TW, (double*)lds, tw);
    if (ok) {
#pragma unroll
        for (int j = 0; j < Q; ++j) B[((int64_t)k * N + reg_elem<P, Q>(p, j)) * H + m] = x[j];
    }
}

// Z, register form: forward, kernel spectrum, inverse; columns in flat (jj, m)
// order.
template <int P, int Q>
__global__ void __launch_bounds__(256) k_fftr_z(FftGeom F, double2* __restrict__ B,
                                                const double* __restrict__ Kc,
                                                const double2* __restrict__ twz) {
    constexpr int N = P * Q, TW = 256 / P;
    extern __shared__ double2 lds[];
    double2* tw = lds + N * TW / 2;                  // after N * TW FP64 values
    const int H = (F.N[0] >> 1) + 1, Ny = F.N[1], nz = F.n[2];
    const int c = threadIdx.x % TW, p = threadIdx.x / TW;
    const int g = blockIdx.x * TW + c;                // flat column over (jj, m)
    const bool ok = g < H * Ny;
    const int jj = ok ? g / H : 0, m = ok ? g - jj * H : 0;
    load_tw(tw, twz, N);
    double2 x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = p + P * q;
        x[q] = (q < Q / 2 && e < nz && ok) ? B[((int64_t)e * Ny + jj) * H + m] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_reg<P, Q, false>(x, p, c, TW, (double*)lds, tw);
    const int ky = min(jj, Ny - jj);
    const int mm = m;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
        const int e = reg_elem<P, Q>(p, j), kz = min(e, N - e);
        const double kv = Kc[((int64_t)kz * ((Ny >> 1) + 1) + ky) * H + mm];
        x[j] = make_double2(x[j].x * kv, x[j].y * kv);
    }
    reg_to_input<P, Q>(x);
    fft_reg<P, Q, true>(x, p, c, TW, (double*)lds, tw);
    if (ok) {
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const int e = reg_elem<P, Q>(p, j);
            if (e < nz) B[((int64_t)e * Ny + jj) * H + m] = x[j];
        }
    }
}

// Y inverse, register form: columns in flat (k, m) order; only j < ny kept.
template <int P, int Q>
__global__ void __launch_bounds__(256) k_fftr_y_inv(FftGeom F, const double2* __restrict__ B,
                                                    double2* __restrict__ A,
                                                    const double2* __restrict__ twy) {
    constexpr int N = P * Q, TW = 256 / P;
    extern __shared__ double2 lds[];
    double2* tw = lds + N * TW / 2;                  // after N * TW FP64 values
    const int H = (F.N[0] >> 1) + 1, ny = F.n[1];
    const int c = threadIdx.x % TW, p = threadIdx.x / TW;
    const int g = blockIdx.x * TW + c;                // flat column over (k, m)
    const bool ok = g < H * F.n[2];
    const int k = ok ? g / H : 0, m = ok ? g - k * H : 0;
    load_tw(tw, twy, N);
    double2 x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        x[q] = ok ? B[((int64_t)k * N + p + P * q) * H + m] : make_double2(0.0, 0.0);
    __syncthreads();
    fft_reg<P, Q, true>(x, p, c, TW, (double*)lds, tw);
    if (ok) {
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const int e = reg_elem<P, Q>(p, j);
            if (e < ny) A[((int64_t)k * ny + e) * H + m] = x[j];
        }
    }
}

// X forward, register form: 256 threads = (256/P rows) x P, p fastest so that
// consecutive threads read consecutive cells of a row.
template <int P, int Q>
__global__ void __launch_bounds__(256) k_fftr_x_fwd(GridDev G, const double* __restrict__ rho,
                                                    int floors, FftGeom F, double2* __restrict__ A,
                                                    const double2* __restrict__ twx) {
    constexpr int N = P * Q, TR = 256 / P;
    extern __shared__ double2 lds[];
    double2* tw = lds + (N + Q) * TR / 2;            // after (N + Q) * TR FP64 values
    const int H = (N >> 1) + 1, nx = F.n[0], rows = F.n[1] * F.n[2];
    const int p = threadIdx.x % P, c = threadIdx.x / P;
    const int row = blockIdx.x * TR + c;
    const int ly = F.logN[1] - 1;                     // ny = Ny / 2
    load_tw(tw, twx, N);
    double2 x[Q];
    const int j = row & ((1 << ly) - 1), k = row >> ly;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = p + P * q;
        double v = 0.0;
        if (q < Q / 2 && e < nx && row < rows) {
            if (rho) {
                v = rho[(int64_t)row * nx + e];
            } else {
                v = G.U[((int64_t)(NG + k) * G.Ny + NG + j) * G.pitch + BSN_XOFF + NG + e];
                if (floors && v < G.rho_floor) v = G.rho_floor;   // src/gpu_kernels.jl:26-46
            }
        }
        x[q] = make_double2(v, 0.0);
    }
    __syncthreads();
    fft_reg<P, Q, false, true>(x, p, c, TR, (double*)lds, tw);
    if (row < rows) {
#pragma unroll
        for (int jx = 0; jx < Q; ++jx) {
            const int e = reg_elem<P, Q>(p, jx);
            if (e < H) A[(int64_t)row * H + e] = x[jx];
        }
    }
}

// X inverse, register form: the full spectrum from its Hermitian half, only
// i < nx kept, scaled, into phi.
template <int P, int Q>
__global__ void __launch_bounds__(256) k_fftr_x_inv(FftGeom F, const double2* __restrict__ A,
                                                    double* __restrict__ phi, double scale,
                                                    const double2* __restrict__ twx) {
    constexpr int N = P * Q, TR = 256 / P;
    extern __shared__ double2 lds[];
    double2* tw = lds + (N + Q) * TR / 2;            // after (N + Q) * TR FP64 values
    const int H = (N >> 1) + 1, nx = F.n[0], rows = F.n[1] * F.n[2];
    const int p = threadIdx.x % P, c = threadIdx.x / P;
    const int row = blockIdx.x * TR + c;
    load_tw(tw, twx, N);
    double2 x[Q];
    const double2* a = A + (int64_t)(row < rows ? row : 0) * H;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = p + P * q;
        double2 v = make_double2(0.0, 0.0);
        if (row < rows) {
            if (e < H) v = a[e];
            else { v = a[N - e]; v.y = -v.y; }
        }
        x[q] = v;
    }
    __syncthreads();
    fft_reg<P, Q, true, true>(x, p, c, TR, (double*)lds, tw);
    if (row < rows) {
#pragma unroll
        for (int jx = 0; jx < Q; ++jx) {
            const int e = reg_elem<P, Q>(p, jx);
            if (e < nx) phi[(int64_t)row * nx + e] = x[jx].x * scale;
        }
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

// Each pass in register form when its axis has N = P*Q with Q in {P, 2P} and
// 16 <= N <= 256, else in the LDS radix-4 form.
enum { PASS_X_FWD, PASS_Y_FWD, PASS_Z, PASS_Y_INV, PASS_X_INV };

struct RegArgs {
    const GridDev* G;
    const double* rho;
    int floors;
    double2* A;
    double2* B;
    const double* Kc;
    const double2* tw;
    double* phi;
    double scale;
};

template <int P, int Q>
static void launch_reg(int pass, hipStream_t s, const FftGeom& F, const RegArgs& a) {
    constexpr int N = P * Q, TW = 256 / P;
    const int H = (F.N[0] >> 1) + 1, rows = F.n[1] * F.n[2];
    const size_t lds = (size_t)(N + Q) * TW * sizeof(double) + (size_t)(N / 2) * sizeof(double2);
    // y and z passes: one flat index over (row, m), rows = nz (y) or Ny (z)
    const int64_t cols = (int64_t)H * (pass == PASS_Z ? F.N[1] : F.n[2]);
    const dim3 gy((unsigned)((cols + TW - 1) / TW));
    const dim3 gx((rows + TW - 1) / TW);
    switch (pass) {
        case PASS_X_FWD:
            hipLaunchKernelGGL((k_fftr_x_fwd<P, Q>), gx, dim3(256), lds, s, *a.G, a.rho, a.floors, F,
                               a.A, a.tw);
            break;
        case PASS_Y_FWD:
            hipLaunchKernelGGL((k_fftr_y_fwd<P, Q>), gy, dim3(256), lds, s, F, a.A, a.B, a.tw);
            break;
        case PASS_Z:
            hipLaunchKernelGGL((k_fftr_z<P, Q>), gy, dim3(256), lds, s, F, a.B, a.Kc, a.tw);
            break;
        case PASS_Y_INV:
            hipLaunchKernelGGL((k_fftr_y_inv<P, Q>), gy, dim3(256), lds, s, F, a.B, a.A, a.tw);
            break;
        default:
            hipLaunchKernelGGL((k_fftr_x_inv<P, Q>), gx, dim3(256), lds, s, F, a.A, a.phi, a.scale,
                               a.tw);
            break;
    }
}

static bool launch_reg_any(int pass, int N, hipStream_t s, const FftGeom& F, const RegArgs& a) {
    switch (N) {
        case 16:   launch_reg<4, 4>(pass, s, F, a); return true;
        case 32:   launch_reg<4, 8>(pass, s, F, a); return true;
        case 64:   launch_reg<8, 8>(pass, s, F, a); return true;
        case 128:  launch_reg<8, 16>(pass, s, F, a); return true;
        case 256:  launch_reg<16, 16>(pass, s, F, a); return true;
        default:   return false;     // longer transforms would hold too many registers
    }
}

hipError_t fft_convolve(hipStream_t s, const FftGeom& F, const GridDev& G, const double* rho,
                        int floors, double2* A, double2* B, const double* Kc,
                        const double2* const tw[3], double* phi, double scale) {
    const int H = (F.N[0] >> 1) + 1;
    const int rows = F.n[1] * F.n[2];
    const int Rx = FFT_CE / F.N[0];
    const int TWy = FFT_CE / F.N[1], TWz = FFT_CE / F.N[2];
    auto lds = [](int N) { return (size_t)(FFT_CE + N / 2) * sizeof(double2); };
    RegArgs a{&G, rho, floors, A, B, Kc, nullptr, phi, scale};
    a.tw = tw[0];
    if (!launch_reg_any(PASS_X_FWD, F.N[0], s, F, a))
        hipLaunchKernelGGL(k_fft_x_fwd, dim3((rows + Rx - 1) / Rx), dim3(256), lds(F.N[0]), s, G, rho,
                           floors, F, A, tw[0]);
    a.tw = tw[1];
    if (!launch_reg_any(PASS_Y_FWD, F.N[1], s, F, a))
        hipLaunchKernelGGL(k_fft_y_fwd, dim3((H + TWy - 1) / TWy, F.n[2]), dim3(256), lds(F.N[1]), s,
                           F, A, B, tw[1]);
    a.tw = tw[2];
    if (!launch_reg_any(PASS_Z, F.N[2], s, F, a))
        hipLaunchKernelGGL(k_fft_z, dim3((H + TWz - 1) / TWz, F.N[1]), dim3(256), lds(F.N[2]), s, F,
                           B, Kc, tw[2]);
    a.tw = tw[1];
    if (!launch_reg_any(PASS_Y_INV, F.N[1], s, F, a))
        hipLaunchKernelGGL(k_fft_y_inv, dim3((H + TWy - 1) / TWy, F.n[2]), dim3(256), lds(F.N[1]), s,
                           F, B, A, tw[1]);
    a.tw = tw[0];
    if (!launch_reg_any(PASS_X_INV, F.N[0], s, F, a))
        hipLaunchKernelGGL(k_fft_x_inv, dim3((rows + Rx - 1) / Rx), dim3(256), lds(F.N[0]), s, F, A,
                           phi, scale, tw[0]);
    return hipGetLastError();
}

}  // namespace bsn
