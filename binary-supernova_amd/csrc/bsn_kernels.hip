// bsn_kernels.hip -- CDNA4 kernels of the Euler right-hand side, the SSP-RK3
// stage update, the CFL reduction, ghost fills and FMR transfer operators.
//
// Kernel map (reference function -> kernel):
//   fill_ghost_3d_outflow!/periodic!  euler3d.jl:33-62        k_ghost_fill
//   _apply_floors_kernel!             gpu_kernels.jl:26-46    k_floors (and inside k_update)
//   _weno_fluxes_{x,y,z}_kernel!      gpu_kernels.jl:51-154   k_flux<DIR>
//   _flux_divergence_kernel!          gpu_kernels.jl:159-169  k_update (fused)
//   _sg_gradient_kernel!              gpu_kernels.jl:291-319  k_update (fused), k_sg_force
//   _bh_gravity_kernel!               gpu_kernels.jl:199-230  update kernels (fused), k_src
//   _sink_sources_kernel!             gpu_kernels.jl:238-284  update kernels (fused), k_src
//   RK3 stage combines                euler3d.jl:244,250,256  k_update (fused), k_ghost_combine
//   _cfl_speeds_kernel! + maximum     gpu_kernels.jl:174-191  k_update (stage 3), k_cfl
//   _prolong_fine_ghosts!/_prolong_all!  fmr3d.jl:206-255     k_prolong
//   _restrict!                        fmr3d.jl:266-285        k_restrict
//   _accumulate_fr_coarse!/_fine!     fmr3d.jl:292-369        k_fr_coarse, k_fr_fine
//   _bc_correct!                      fmr3d.jl:379-422        k_bc_correct
//
// All work is described by device-resident tile lists (struct Tile) or by
// per-grid prefix offsets, so one launch sweeps any number of grids of any
// shape.  Tile-list kernels are persistent: a workgroup takes a contiguous
// chunk of the list, which keeps neighbouring tiles (shared halo cells) on one
// CU and one XCD L2.
#include <hip/hip_runtime.h>
#include <mutex>
#include "bsn_internal.hpp"
#include "bsn_physics.hpp"
#include "bsn_launch.hpp"

namespace bsn {

#define NG BSN_NG

__device__ __forceinline__ int64_t uidx(const GridDev& G, int i, int j, int k) {
    return ((int64_t)k * G.Ny + j) * G.pitch + BSN_XOFF + i;
}

// Global-address-space view of a device pointer.  The state and flux
// pointers come out of the GridDev table in memory, so the compiler cannot
// infer their address space and would emit flat loads and stores; flat
// operations also count against the scalar-load counter, and every scalar
// reload of a table field would then wait for all HBM loads in flight.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// base + a 32-bit byte offset.  With a uniform base (scalar registers) and a
// per-lane offset, a store takes the global saddr form and needs no 64-bit
// address arithmetic.  Callers keep the uniform parts (the field, the z plane)
// in the base and only the in-plane part in the offset, so the offset stays
// far below 4 GiB.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* at_off(T* base, uint32_t off) {
    return (__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) char*)base + off);
}

// Load through a raw buffer descriptor: a uniform base (scalar registers, a
// range of 2 GiB) and a 32-bit per-lane byte offset.  The compiler folds
// a plain pointer's per-lane offset into the base and then carries that
// 64-bit per-lane address from one loop iteration to the next, one 64-bit
// VALU addition per load; through the descriptor the base stays scalar.
// The base must be uniform over the wave.  NT: streaming (nontemporal).
template <typename T, bool NT = false>
__device__ __forceinline__ T bld(const T* base, uint32_t off) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte loads");
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void*)base, 0, 0x7fffffff, 0x00020000);
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, NT ? 2 : 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, NT ? 2 : 0));
}

__device__ __forceinline__ void chunk(int ntiles, int& t0, int& t1) {
    t0 = (int)(((int64_t)blockIdx.x * ntiles) / gridDim.x);
    t1 = (int)(((int64_t)(blockIdx.x + 1) * ntiles) / gridDim.x);
}

// Lane -> (x, row) of a column tile: 64 lanes along x and one row per wave,
// or, for a half-width tile (COL_HALF in a), 32 lanes along x and two rows per
// wave.  Uniform over the workgroup.
__device__ __forceinline__ void col_lane(int a, int b, int& i, int& row) {
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    if (a & COL_HALF) { i = (a & ~COL_HALF) + (lx & 31); row = b + 2 * ly + (lx >> 5); }
    else              { i = a + lx; row = b + ly; }
}

__device__ __forceinline__ int find_grid(const int64_t* __restrict__ off, int ngrids, int64_t e) {
    int lo = 0, hi = ngrids - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Ghost cells.  The set of ghost cells of a grid is split into three disjoint
// slabs: x-slabs (ghost i, all j, k), y-slabs (active i, ghost j, all k) and
// z-slabs (active i, j, ghost k) -- the same partition as the reference's
// fill order (euler3d.jl:36-43, fmr3d.jl:228-254).

__device__ __forceinline__ int64_t ghost_count(const GridDev& G) {
    return 2 * NG * ((int64_t)G.Ny * G.Nz + (int64_t)G.nx * G.Nz + (int64_t)G.nx * G.ny);
}

__device__ __forceinline__ void ghost_decode(const GridDev& G, int64_t t, int& i, int& j, int& k) {
    const int64_t A = 2 * NG * (int64_t)G.Ny * G.Nz;
    const int64_t B = 2 * NG * (int64_t)G.nx * G.Nz;
    if (t < A) {
        int ig = (int)(t % (2 * NG)); t /= 2 * NG;
        j = (int)(t % G.Ny); k = (int)(t / G.Ny);
        i = ig < NG ? ig : G.nx + ig;
    } else if (t < A + B) {
        t -= A;
        i = NG + (int)(t % G.nx); t /= G.nx;
        int jg = (int)(t % (2 * NG)); k = (int)(t / (2 * NG));
        j = jg < NG ? jg : G.ny + jg;
    } else {
        t -= A + B;
        i = NG + (int)(t % G.nx); t /= G.nx;
        j = NG + (int)(t % G.ny);
        int kg = (int)(t / G.ny);
        k = kg < NG ? kg : G.nz + kg;
    }
}

__device__ __forceinline__ int bc_src(int a, int n, bool periodic) {
    if (periodic) return a < NG ? a + n : (a >= NG + n ? a - n : a);
    return a < NG ? NG : (a >= NG + n ? NG + n - 1 : a);
}

// Ghost fill by boundary condition.  A ghost cell takes the value of the
// active cell its index maps to per axis; this equals the reference's
// x -> y -> z sequence of slab copies (each slab copy only moves values that
// an earlier copy placed).  With copy_un, the raw value of every ghost cell
// is first saved to Un (the RK3 base state is a copy of the input).
__global__ void k_ghost_fill(const GridDev* __restrict__ grids, const int64_t* __restrict__ off,
                             int ngrids, int copy_un) {
    const int64_t total = off[ngrids];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int g = find_grid(off, ngrids, e);
        const GridDev& G = grids[g];
        int i, j, k;
        ghost_decode(G, e - off[g], i, j, k);
        const int si = bc_src(i, G.nx, G.bc & 1), sj = bc_src(j, G.ny, G.bc & 2),
                  sk = bc_src(k, G.nz, G.bc & 4);
        const int64_t d = uidx(G, i, j, k), s = uidx(G, si, sj, sk);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            if (copy_un) G.Un[q * G.fs + d] = G.U[q * G.fs + d];
            G.U[q * G.fs + d] = G.U[q * G.fs + s];
        }
    }
}

// Stage-3 combination on ghost cells, where the reference's dU is zero
// (euler3d.jl:138 fills dU with zeros; the combine runs over the whole array).
// With ghost mapping the stage input's ghost value is read from the active
// cell it maps to.
__global__ void k_ghost_combine(const GridDev* __restrict__ grids, const int64_t* __restrict__ off,
                                int ngrids, int stage) {
    const int64_t total = off[ngrids];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int g = find_grid(off, ngrids, e);
        const GridDev& G = grids[g];
        int i, j, k;
        ghost_decode(G, e - off[g], i, j, k);
        const int64_t d = uidx(G, i, j, k);
        const int64_t su = (G.lmode & 1) ? uidx(G, bc_src(i, G.nx, G.bc & 1), bc_src(j, G.ny, G.bc & 2),
                                                bc_src(k, G.nz, G.bc & 4))
                                         : d;
        const double dt = G.scal[SCAL_DT];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const double un = G.Un[q * G.fs + d], u = G.U[q * G.fs + su];
            double v;
            if (stage == 1)      v = un + dt * 0.0;
            else if (stage == 2) v = 0.75 * un + 0.25 * u + (0.25 * dt) * 0.0;
            else                 v = (1.0 / 3.0) * un + (2.0 / 3.0) * u + ((2.0 / 3.0) * dt) * 0.0;
            G.Uout[q * G.fs + d] = v;
        }
    }
}

// a / b: IEEE division for the FP64 form; for the split form (FAST) the
// product with a reciprocal good to a few ulp, which its FP64 state needs for
// range and accuracy, not for bit parity.
template <bool FAST>
__device__ __forceinline__ double ddiv(double a, double b) {
    if (FAST) return a * sp::rcp64(b);
    return a / b;
}

// src/gpu_kernels.jl:26-46 on the five values of one cell.
template <bool FAST = false>
__device__ __forceinline__ bool apply_floor(double (&u)[5], double rho_floor, double P_floor,
                                            double g) {
    const double r = u[0];
    if (r < rho_floor) {
        u[0] = rho_floor; u[1] = 0.0; u[2] = 0.0; u[3] = 0.0;
        u[4] = P_floor / (g - 1.0);
        return true;
    }
    const double KE = ddiv<FAST>(0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]), r);
    const double P = (g - 1.0) * (u[4] - KE);
    if (P < P_floor) { u[4] = P_floor / (g - 1.0) + KE; return true; }
    return false;
}

// Loads the five values a stage sees at padded cell (i, j, k) of its input
// U, where only axis AX may be a ghost index.  With lmode bit 0 a ghost cell
// reads the active cell its index maps to (the value the reference's ghost
// fill would have copied there, unfloored: the fill precedes the floors,
// euler3d.jl:189-196); with lmode bit 1 an active cell is floored on load.
// NT: streaming (nontemporal) loads, for callers that read each cell once.
template <int AX, bool FLOOR = true, bool FAST = false, bool NT = false>
__device__ __forceinline__ void load_cell(const GridDev& G, int i, int j, int k, double (&u)[5]) {
    int si = i, sj = j, sk = k;
    if (G.lmode & 1) {
        if (AX == 0) si = bc_src(i, G.nx, G.bc & 1);
        if (AX == 1) sj = bc_src(j, G.ny, G.bc & 2);
        if (AX == 2) sk = bc_src(k, G.nz, G.bc & 4);
    }
    const int64_t o = uidx(G, si, sj, sk);
#pragma unroll
    for (int q = 0; q < 5; ++q)
        u[q] = NT ? __builtin_nontemporal_load(&gmem(G.U)[q * G.fs + o]) : gmem(G.U)[q * G.fs + o];
    // only active cells are floored (callers pass active cross indices; a
    // ghost cell holds the raw value of a ghost fill or a prolongation)
    const bool active = AX == 0 ? (i >= NG && i < NG + G.nx)
                      : AX == 1 ? (j >= NG && j < NG + G.ny) : (k >= NG && k < NG + G.nz);
    if (FLOOR && (G.lmode & 2) && active) apply_floor<FAST>(u, G.rho_floor, G.P_floor, G.gamma);
}

// Floors on active cells (apply_floors_3d!, euler3d.jl:73-80); with copy_un
// the raw value is saved to Un first.
__global__ void k_floors(const GridDev* __restrict__ grids, const int64_t* __restrict__ off,
                         int ngrids, int copy_un) {
    const int64_t total = off[ngrids];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int g = find_grid(off, ngrids, e);
        const GridDev& G = grids[g];
        int64_t t = e - off[g];
        const int i = NG + (int)(t % G.nx); t /= G.nx;
        const int j = NG + (int)(t % G.ny);
        const int k = NG + (int)(t / G.ny);
        const int64_t c = uidx(G, i, j, k);
        double u[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) u[q] = G.U[q * G.fs + c];
        if (copy_un) {
#pragma unroll
            for (int q = 0; q < 5; ++q) G.Un[q * G.fs + c] = u[q];
        }
        // the split form floors with the same division wherever it floors
        const bool fired = G.prec == 1 ? apply_floor<true>(u, G.rho_floor, G.P_floor, G.gamma)
                                       : apply_floor<false>(u, G.rho_floor, G.P_floor, G.gamma);
        if (fired) {
#pragma unroll
            for (int q = 0; q < 5; ++q) G.U[q * G.fs + c] = u[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Interface fluxes.  One workgroup stages a pencil tile of cells (five
// conserved values reduced to rho, m, P) in LDS, six cells deep along the
// sweep direction, then each thread computes one interface.
//
//   DIR 0 (x): tile (g, f0, j0, k): interfaces f0..f0+W-1 (f < nx) along x
//              for rows j0..j0+H-1.  Lanes run along x.  A tile with
//              a = -1 instead computes the last interface f = nx for the
//              rows j0..j0+W*H-1, gathering from HBM.  Each wave stages and
//              computes its own rows (64 / W of them), with no block
//              barrier.
//   DIR 1 (y): tile (g, f0, i0, k): interfaces f0..f0+H-1 (f <= ny) along y
//              for cells i0..i0+W-1.  Lanes run along x (coalesced).
//   DIR 2 (z): tile (g, f0, i0, j): as DIR 1 along z.

template <typename T> struct Prec;
template <> struct Prec<double> {
    template <int DIR>
    __device__ static void flux(const double (&c)[4][6], const double (&p)[6], const GridDev& G,
                                double F[5]) {
        ex::interface_flux<DIR>(c, p, G.gamma, G.gfac, F);
    }
    __device__ static double pressure(const double (&u)[5], double g) {
        return ex::cell_pressure(u[0], u[1], u[2], u[3], u[4], g);
    }
};
template <> struct Prec<float> {
    template <int DIR>
    __device__ static void flux(const float (&c)[4][6], const float (&p)[6], const GridDev& G,
                                float F[5]) {
        sp::interface_flux<DIR>(c, p, G.gf32[0], G.gf32[1], G.gf32[2], F);
    }
    __device__ static double pressure(const double (&u)[5], double g) {
        return sp::cell_pressure(u[0], u[1], u[2], u[3], u[4], g);
    }
};

// Stores the flux of interface f along DIR with cross indices ci, cj
// (active, 0-based; ci runs along x for DIR 1 and 2, along y for DIR 0).
// Streaming stores: the update reads each face flux once, from HBM (1.2% off
// the split step at 256^3 against cached stores).
template <int DIR, typename T>
__device__ __forceinline__ void store_face(const GridDev& G, const T (&F)[5], int f, int ci, int cj) {
    auto Fp = gmem((T*)G.F[DIR]);
    int64_t o;
    if (DIR == 0)      o = ((int64_t)cj * G.ny + ci) * G.fp[0] + f;
    else if (DIR == 1) o = ((int64_t)cj * (G.ny + 1) + f) * G.fp[1] + ci;
    else               o = ((int64_t)f * G.ny + cj) * G.fp[2] + ci;
#pragma unroll
    for (int q = 0; q < 5; ++q) __builtin_nontemporal_store(F[q], &Fp[q * G.ffs[DIR] + o]);
}

template <int DIR, typename T, int W, int H>
__global__ __launch_bounds__(W * H, (DIR == 0 && W == 64 && H == 4 && sizeof(T) == 4) ? 6 : 1) void k_flux(const GridDev* __restrict__ grids,
                                                 const Tile* __restrict__ tiles, int ntiles) {
    constexpr int NT = W * H;
    constexpr bool FAST = sizeof(T) == 4;        // split form
    constexpr int SA = DIR == 0 ? H : H + 5;     // LDS rows
    constexpr int SB = DIR == 0 ? W + 5 : W;     // LDS row length
    // x tiles of 256 threads, 64, 32 or 16 wide: each wave stages and
    // computes its own rows, with no block barrier.  The split 64 x 4 tile is
    // held to 80 VGPRs, 6 waves per SIMD (a few values spilled): 0.367 ms
    // against 0.376 at 91 VGPRs, 5 waves.
    constexpr bool WAVE = DIR == 0 && W * H == 256 && (W == 64 || W == 32 || W == 16);
    __shared__ T lds[5][SA][SB];
    const int tid = threadIdx.x;
    int t0, t1;
    chunk(ntiles, t0, t1);
    // interface f along DIR with cross indices ci, cj (active, 0-based)
    auto emit = [&](const GridDev& G, const T (&c)[4][6], const T (&p)[6], int f, int ci, int cj) {
        T F[5];
        Prec<T>::template flux<DIR>(c, p, G, F);
        store_face<DIR>(G, F, f, ci, cj);
    };
    for (int t = t0; t < t1; ++t) {
        const Tile tl = tiles[t];
        const GridDev& G = grids[tl.g];
        const int Nx = G.nx + 2 * NG;
        T c[4][6], p[6];
        if (tl.a < 0) {
            // edge tile (x only): the last interface f = nx of NT consecutive
            // rows from b, gathered from HBM.  The tile is uniform, so every
            // thread of the block takes this branch.
            const int ci = tl.b + tid, cj = tl.c;
            const int f = DIR == 0 ? G.nx : DIR == 1 ? G.ny : G.nz;
            if (ci < (DIR == 0 ? G.ny : G.nx)) {
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    double u[5];
                    if (DIR == 0)      load_cell<0, true, FAST>(G, f + m, NG + ci, NG + cj, u);
                    else if (DIR == 1) load_cell<1, true, FAST>(G, NG + ci, f + m, NG + cj, u);
                    else               load_cell<2, true, FAST>(G, NG + ci, NG + cj, f + m, u);
                    const double pr = Prec<T>::pressure(u, G.gamma);
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[q][m] = (T)u[q];
                    p[m] = (T)pr;
                }
                emit(G, c, p, f, ci, cj);
            }
            continue;
        }
        if constexpr (WAVE) {
            // each wave stages and computes its own RW = 64 / W rows of the
            // tile (lane -> row hb + lane / W, x = lane % W): cells f0 .. f0+W+4
            // of each row, both loads issued before either is used (indices
            // clamped into the grid), then floors and pressures.  The wave's
            // first row is uniform: its base stays in scalar registers, and a
            // lane's row past it goes into its 32-bit load and store offsets.
            constexpr int RW = 64 / W;
            const int hb = __builtin_amdgcn_readfirstlane(tid >> 6) * RW, lane = tid & 63;
            const int x = lane % W, h = hb + lane / W;
            const bool row_ok = tl.b + h < G.ny;
            const int pjb = NG + min(tl.b + hb, G.ny - 1), pk = NG + tl.c;
            const int pj = NG + min(tl.b + h, G.ny - 1);
            const double* rowU = G.U + uidx(G, 0, pjb, pk);
            const uint32_t roff = (uint32_t)((pj - pjb) * G.pitch) * 8u;
            double u[2][5];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                int pi = min(tl.a + (r ? W + min(x, 4) : x), Nx - 1);
                if (G.lmode & 1) pi = bc_src(pi, G.nx, G.bc & 1);
#pragma unroll
                for (int q = 0; q < 5; ++q) u[r][q] = bld(rowU + q * G.fs, roff + (uint32_t)pi * 8u);
            }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int b = r ? W + x : x, pi = tl.a + b;
                if ((r == 0 || x < 5) && row_ok && pi < Nx) {
                    if ((G.lmode & 2) && pi >= NG && pi < NG + G.nx)
                        apply_floor<FAST>(u[r], G.rho_floor, G.P_floor, G.gamma);
                    const double pr = Prec<T>::pressure(u[r], G.gamma);
#pragma unroll
                    for (int q = 0; q < 4; ++q) lds[q][h][b] = (T)u[r][q];
                    lds[4][h][b] = (T)pr;
                }
            }
            // the rows' cells come from other lanes of this wave: the LDS
            // runs one wave's operations in order, and the fences keep the
            // compiler from moving the reads above the writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int f = tl.a + x, ci = tl.b + h;
            if (f < G.nx && ci < G.ny) {
#pragma unroll
                for (int m = 0; m < 6; ++m) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[q][m] = lds[q][h][x + m];
                    p[m] = lds[4][h][x + m];
                }
                T F[5];
                Prec<T>::template flux<0>(c, p, G, F);
                T* rowF = (T*)G.F[0] + ((int64_t)tl.c * G.ny + tl.b + hb) * G.fp[0];
                const uint32_t fo = (uint32_t)((h - hb) * G.fp[0] + f) * (uint32_t)sizeof(T);
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    __builtin_nontemporal_store(F[q], at_off(rowF + q * G.ffs[0], fo));
            }
            // the next tile's stores follow this tile's reads (same order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            continue;
        }
        // stage cells into LDS
        for (int e = tid; e < SA * SB; e += NT) {
            const int a = e / SB, b = e - a * SB;
            int pi, pj, pk;
            bool ok;
            if (DIR == 0) {
                pi = tl.a + b; pj = NG + tl.b + a; pk = NG + tl.c;
                ok = pi < Nx && tl.b + a < G.ny;
            } else if (DIR == 1) {
                pi = NG + tl.b + b; pj = tl.a + a; pk = NG + tl.c;
                ok = tl.b + b < G.nx && pj < G.ny + 2 * NG;
            } else {
                pi = NG + tl.b + b; pj = NG + tl.c; pk = tl.a + a;
                ok = tl.b + b < G.nx && pk < G.nz + 2 * NG;
            }
            if (ok) {
                double u[5];
                load_cell<DIR, true, FAST>(G, pi, pj, pk, u);
                const double pr = Prec<T>::pressure(u, G.gamma);
#pragma unroll
                for (int q = 0; q < 4; ++q) lds[q][a][b] = (T)u[q];
                lds[4][a][b] = (T)pr;
            }
        }
        __syncthreads();
        const int x = tid % W, h = tid / W;
        int f, ci;
        bool valid;
        if (DIR == 0) {
            f = tl.a + x; ci = tl.b + h;
            valid = f < G.nx && ci < G.ny;
        } else {
            f = tl.a + h; ci = tl.b + x;
            valid = f <= (DIR == 1 ? G.ny : G.nz) && ci < G.nx;
        }
        if (valid) {
#pragma unroll
            for (int m = 0; m < 6; ++m) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    c[q][m] = DIR == 0 ? lds[q][h][x + m] : lds[q][h + m][x];
                p[m] = DIR == 0 ? lds[4][h][x + m] : lds[4][h + m][x];
            }
            emit(G, c, p, f, ci, tl.c);
        }
        __syncthreads();
    }
}

// x interface fluxes as short register walks.  Tile (g, f0, j0, k | n): a
// block of four waves stages rows j0 .. j0+3 of plane k, cells f0 .. f0+n+4
// (padded x), in LDS; lane l of wave w walks row j0+w through its share of
// the n <= 64 KMAX interfaces f0 .. f0+n-1 (KMAX consecutive ones when the
// tile is full) with a six-cell register window read from LDS, and the split
// form carries the smoothness indicators that consecutive interfaces share.
// With OUT the fluxes go back through the row's LDS slots, so that the lanes
// store them to HBM contiguously; without it each lane stores its fluxes
// as they are formed (cached stores: the L2 joins the lanes' strided words
// into whole lines), which needs no flux registers.  Cell b of a row sits at b + b/4: a lane's
// cells start five words after its neighbour's, so the lanes of a wave read
// distinct banks.  Interfaces 0 .. nx-1 are covered by such tiles; an edge
// tile (g, -1, r0, 0) computes the last interface f = nx of the 256 rows
// r0 .. r0+255 (row r = k ny + j), one per thread, gathering from HBM.
template <typename T, int KMAX, bool OUT>
__global__ __launch_bounds__(256) void k_fxwalk(const GridDev* __restrict__ grids,
                                                const Tile* __restrict__ tiles, int ntiles) {
    constexpr bool FAST = sizeof(T) == 4;        // split form
    constexpr int NC = 64 * KMAX + 5;            // staged cells per row
    constexpr int SB = NC + NC / 4 + 1;          // LDS row length after the skew
    __shared__ T lds[5][4][SB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    auto at = [](int b) { return b + (b >> 2); };
    int t0, t1;
    chunk(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) {
        const Tile tl = tiles[t];
        const GridDev& G = grids[tl.g];
        if (tl.a < 0) {
            // edge tile: uniform over the block, so no thread reaches the
            // barriers below while another takes this branch
            const int r = tl.b + tid;
            if (r < G.ny * G.nz) {
                const int jr = r % G.ny, kr = r / G.ny;
                T c[4][6], p[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    double u[5];
                    load_cell<0, true, FAST>(G, G.nx + m, NG + jr, NG + kr, u);
                    const double pr = Prec<T>::pressure(u, G.gamma);
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[q][m] = (T)u[q];
                    p[m] = (T)pr;
                }
                T F[5];
                Prec<T>::template flux<0>(c, p, G, F);
                store_face<0>(G, F, G.nx, jr, kr);
            }
            continue;
        }
        const int f0 = tl.a, j0 = tl.b, k = tl.c & ZT_K0_MASK, n = tl.c >> ZT_LEN_SHIFT;
        const int nc = n + 5;
        {
            // every thread issues all of its loads before it uses the first
            // (indices clamped into the grid, so the loads need no branch),
            // so that the tile costs one memory round trip, not NE; the floors
            // follow the loads
            constexpr int NE = (4 * NC + 255) / 256;
            double u[NE][5];
#pragma unroll
            for (int r = 0; r < NE; ++r) {
                const int e = min(tid + 256 * r, 4 * nc - 1);
                const int a = e / nc, b = e - a * nc;
                load_cell<0, false, FAST>(G, f0 + b, NG + min(j0 + a, G.ny - 1), NG + k, u[r]);
            }
#pragma unroll
            for (int r = 0; r < NE; ++r) {
                const int e = tid + 256 * r;
                const int a = e / nc, b = e - a * nc;
                if (e < 4 * nc && j0 + a < G.ny) {
                    if ((G.lmode & 2) && f0 + b >= NG && f0 + b < NG + G.nx)
                        apply_floor<FAST>(u[r], G.rho_floor, G.P_floor, G.gamma);
                    const double pr = Prec<T>::pressure(u[r], G.gamma);
#pragma unroll
                    for (int q = 0; q < 4; ++q) lds[q][a][at(b)] = (T)u[r][q];
                    lds[4][a][at(b)] = (T)pr;
                }
            }
        }
        __syncthreads();
        const int j = j0 + w;
        const int b0 = (lane * n) >> 6, len = (((lane + 1) * n) >> 6) - b0;
        T Fb[OUT ? KMAX : 1][5];
        if (j < G.ny && len > 0) {
            T c[4][6], p[6];
            auto place = [&](int slot, int b) {
#pragma unroll
                for (int q = 0; q < 4; ++q) c[q][slot] = lds[q][w][at(b)];
                p[slot] = lds[4][w][at(b)];
            };
#pragma unroll
            for (int m = 0; m < 5; ++m) place(m, b0 + m);
            sp::WenoCarry wc[5];
            if constexpr (FAST) sp::interface_walk_start(c, p, wc);
            auto step = [&](int m, T (&F)[5]) {
                place(5, b0 + m + 5);
                if constexpr (FAST)
                    sp::interface_flux<0, true>(c, p, G.gf32[0], G.gf32[1], G.gf32[2], F, wc);
                else
                    Prec<T>::template flux<0>(c, p, G, F);
#pragma unroll
                for (int s = 0; s < 5; ++s) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[q][s] = c[q][s + 1];
                    p[s] = p[s + 1];
                }
            };
            if constexpr (OUT) {
                // unrolled, so that the flux registers are indexed statically
#pragma unroll
                for (int m = 0; m < KMAX; ++m)
                    if (m < len) step(m, Fb[m]);
            } else {
                auto Fp = gmem((T*)G.F[0]);
                const int64_t o = ((int64_t)k * G.ny + j) * G.fp[0] + f0 + b0;
                for (int m = 0; m < len; ++m) {
                    step(m, Fb[0]);
#pragma unroll
                    for (int q = 0; q < 5; ++q) Fp[q * G.ffs[0] + o + m] = Fb[0][q];
                }
            }
        }
        if (!OUT) {
            __syncthreads();
            continue;
        }
        // every lane of the block has read its cells: the fluxes replace them
        __syncthreads();
        if (j < G.ny) {
#pragma unroll
            for (int m = 0; m < KMAX; ++m)
                if (m < len) {
#pragma unroll
                    for (int q = 0; q < 5; ++q) lds[q][w][at(b0 + m)] = Fb[OUT ? m : 0][q];
                }
        }
        __syncthreads();
        if (j < G.ny) {
            auto Fp = gmem((T*)G.F[0]);
            const int64_t row = ((int64_t)k * G.ny + j) * G.fp[0] + f0;
            // streaming stores, as for the other sweeps
            for (int e = lane; e < n; e += 64) {
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    __builtin_nontemporal_store(lds[q][w][at(e)], &Fp[q * G.ffs[0] + row + e]);
            }
        }
        __syncthreads();
    }
}

// y (DIR 1) and z (DIR 2) interface fluxes as register walks.  Tile
// (g, i0, c0, f0 | len): 64 x 4 threads over (i, c) with c = k for DIR 1 and
// c = j for DIR 2; each thread walks its column along DIR through the
// interfaces f0 .. f0+len-1 with a six-cell register window, loading one new
// cell per interface straight from HBM (lanes along x: coalesced), and the
// split form carries the smoothness indicators that consecutive interfaces
// share.  A run of len interfaces reads len + 5 cells.
template <int DIR, typename T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1) void k_fwalk(const GridDev* __restrict__ grids,
                                               const Tile* __restrict__ tiles, int ntiles) {
    constexpr bool FAST = sizeof(T) == 4;        // split form
    int t0, t1;
    chunk(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) {
        const Tile tl = tiles[t];
        const GridDev& G = grids[tl.g];
        const int f0 = tl.c & ZT_K0_MASK, len = tl.c >> ZT_LEN_SHIFT;
        int i, cc;
        col_lane(tl.a, tl.b, i, cc);
        // the wave's first row (uniform) and this lane's row past it (0, or 1
        // in the second half of a half-width tile): the walk's addresses are
        // a scalar base per cell or face and a lane offset fixed for the tile
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int cc0 = (tl.a & COL_HALF) ? tl.b + 2 * wv : tl.b + wv;
        const int dr = cc - cc0;
        if (i >= G.nx || cc >= (DIR == 1 ? G.nz : G.ny)) continue;
        const uint32_t uoff =
            (uint32_t)(((DIR == 1 ? (int64_t)G.Ny * G.pitch : (int64_t)G.pitch) * dr + NG + i) * 8);
        const uint32_t foff =
            (uint32_t)(((DIR == 1 ? (int64_t)(G.ny + 1) * G.fp[1] : (int64_t)G.fp[2]) * dr + i) * sizeof(T));
        T c[4][6], p[6];
        double nxt[5];
        int npf = 0;
        // cells along the walk are read once (halo cells of neighbouring
        // runs aside): streaming loads.  The load is raw and the floor is
        // applied when the cell is placed, so that the next cell's load stays
        // in flight during the current interface's flux.  A ghost index along
        // the walk reads the active cell it maps to, as load_cell does.
        auto fetch = [&](int pf) {
            npf = pf;
            int sp_ = pf;
            if (G.lmode & 1) sp_ = bc_src(pf, DIR == 1 ? G.ny : G.nz, G.bc & (DIR == 1 ? 2 : 4));
            const int64_t base = DIR == 1 ? ((int64_t)(NG + cc0) * G.Ny + sp_) * G.pitch + BSN_XOFF
                                          : ((int64_t)sp_ * G.Ny + NG + cc0) * G.pitch + BSN_XOFF;
#pragma unroll
            for (int q = 0; q < 5; ++q) nxt[q] = bld<double, true>(G.U + q * G.fs + base, uoff);
        };
        auto place = [&](int slot) {
            if ((G.lmode & 2) && npf >= NG && npf < NG + (DIR == 1 ? G.ny : G.nz))
                apply_floor<FAST>(nxt, G.rho_floor, G.P_floor, G.gamma);
            const double pr = Prec<T>::pressure(nxt, G.gamma);
#pragma unroll
            for (int q = 0; q < 4; ++q) c[q][slot] = (T)nxt[q];
            p[slot] = (T)pr;
        };
#pragma unroll
        for (int m = 0; m < 5; ++m) { fetch(f0 + m); place(m); }
        fetch(f0 + 5);
        sp::WenoCarry wc[5];
        if constexpr (FAST) sp::interface_walk_start(c, p, wc);
        const int fend = f0 + len;
        auto step = [&](int f) {
            place(5);                                   // cell f+5 (padded)
            if (f + 1 < fend) fetch(f + 6);             // in flight during the flux
            T F[5];
            if constexpr (FAST)
                sp::interface_flux<DIR, true>(c, p, G.gf32[0], G.gf32[1], G.gf32[2], F, wc);
            else
                Prec<T>::template flux<DIR>(c, p, G, F);
            // store_face's layout, from the wave's first row
            T* fb = (T*)G.F[DIR] + (DIR == 1 ? ((int64_t)cc0 * (G.ny + 1) + f) * G.fp[1]
                                             : ((int64_t)f * G.ny + cc0) * G.fp[2]);
#pragma unroll
            for (int q = 0; q < 5; ++q) __builtin_nontemporal_store(F[q], at_off(fb + q * G.ffs[DIR], foff));
#pragma unroll
            for (int m = 0; m < 5; ++m) {
#pragma unroll
                for (int q = 0; q < 4; ++q) c[q][m] = c[q][m + 1];
                p[m] = p[m + 1];
            }
        };
        int f = f0;
        if constexpr (FAST) {
            // six interfaces per pass: the window's slots come back to the
            // same registers after six shifts, so the shifts cost no moves
            // (split form: 754 -> 695 VALU instructions per interface; the
            // FP64 form would double its registers).  The split walk is held
            // to 128 VGPRs, 4 waves per SIMD, with a few values spilled: the
            // y sweep took 0.295 ms against 0.305 at 133 VGPRs and 3 waves.
            for (; f + 6 <= fend; f += 6) {
#pragma unroll
                for (int u = 0; u < 6; ++u) step(f + u);
            }
        }
        for (; f < fend; ++f) step(f);
    }
}

// ---------------------------------------------------------------------------
// Fused flux divergence + phi source + SSP-RK3 stage combine + floors
// (+ the CFL wave-speed maximum of the new state on stage 3).
// Tile (g, i0, j0, k0): 64 x 4 threads over (i, j); each thread walks KZ
// cells along z and carries the z-face flux.  stage 0 stores dU instead
// (euler3d_rhs!).

template <bool FAST = false>
__device__ __forceinline__ uint64_t cfl_speed_bits(const double (&u)[5], const GridDev& G) {
    // src/gpu_kernels.jl:180-189
    const double g = G.gamma;
    const double r = jl_max(u[0], 1e-30);
    const double vx = ddiv<FAST>(u[1], r), vy = ddiv<FAST>(u[2], r), vz = ddiv<FAST>(u[3], r);
    const double KE = 0.5 * (vx * vx + vy * vy + vz * vz);
    const double P = jl_max(((g - 1.0) * (ddiv<FAST>(u[4], r) - KE)) * r, 0.0);
    const double cs = sqrt(ddiv<FAST>(g * P, r));
    const double s = jl_max(jl_max((fabs(vx) + cs) * G.idx, (fabs(vy) + cs) * G.idy),
                            (fabs(vz) + cs) * G.idz);
    return (uint64_t)__double_as_longlong(s);
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t w = (uint64_t)__shfl_xor((unsigned long long)v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__device__ __forceinline__ void block_max_to(uint64_t v, unsigned long long* dst) {
    __shared__ uint64_t red[16];
    v = wave_max_u64(v);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
        if (m) atomicMax(dst, (unsigned long long)m);
    }
    __syncthreads();
}

// Point-mass gravity on active cell (i, j, k) (0-based) with stage-input
// state u: d += (0, rho a, rho v.a) (src/gpu_kernels.jl:199-230).
__device__ __forceinline__ void bh_gravity_add(const SrcDev& S, int i, int j, int k,
                                               const double (&u)[5], double (&d)[5]) {
    const double xc = S.gx0[0] + ((double)(i + 1) - 0.5) * S.dx;
    const double yc = S.gx0[1] + ((double)(j + 1) - 0.5) * S.dy;
    const double zc = S.gx0[2] + ((double)(k + 1) - 0.5) * S.dz;
    const double r = u[0];
    const double vx = u[1] / r, vy = u[2] / r, vz = u[3] / r;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int n = 0; n < S.nbh; ++n) {
        const double ddx = xc - S.px[n], ddy = yc - S.py[n], ddz = zc - S.pz[n];
        const double r2 = ddx * ddx + ddy * ddy + ddz * ddz + S.eps[n] * S.eps[n];
        const double r3 = r2 * sqrt(r2);
        const double fac = -S.mass[n] / r3;
        ax += fac * ddx;
        ay += fac * ddy;
        az += fac * ddz;
    }
    d[1] += r * ax;
    d[2] += r * ay;
    d[3] += r * az;
    d[4] += r * (vx * ax + vy * ay + vz * az);
}

// Gas sinks on active cell (i, j, k): torque-free or uniform drain inside
// each sink radius (src/gpu_kernels.jl:238-284).
__device__ __forceinline__ void sink_add(const SrcDev& S, int i, int j, int k,
                                         const double (&u)[5], double (&d)[5]) {
    const double r = u[0], mx = u[1], my = u[2], mz = u[3], E = u[4];
    if (!(r > S.rho_sink_min)) return;
    const double xc = S.sx0[0] + ((double)(i + 1) - 0.5) * S.dx;
    const double yc = S.sx0[1] + ((double)(j + 1) - 0.5) * S.dy;
    const double zc = S.sx0[2] + ((double)(k + 1) - 0.5) * S.dz;
    for (int n = 0; n < S.nsink; ++n) {
        const double ddx = xc - S.spx[n], ddy = yc - S.spy[n], ddz = zc - S.spz[n];
        const double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        if (!(dist < S.rs[n])) continue;
        const double t = S.ts[n];
        d[0] -= r / t;
        if (S.torque_free) {
            const double rx = ddx / dist, ry = ddy / dist, rz = ddz / dist;
            const double vx = mx / r, vy = my / r, vz = mz / r;
            const double vrx = vx - S.svx[n], vry = vy - S.svy[n], vrz = vz - S.svz[n];
            const double v_r = vrx * rx + vry * ry + vrz * rz;
            const double e_int = E / r - 0.5 * (vx * vx + vy * vy + vz * vz);
            d[1] -= r * v_r * rx / t;
            d[2] -= r * v_r * ry / t;
            d[3] -= r * v_r * rz / t;
            d[4] -= (0.5 * r * (v_r * v_r) + r * e_int) / t;
        } else {
            d[1] -= mx / t;
            d[2] -= my / t;
            d[3] -= mz / t;
            d[4] -= E / t;
        }
    }
}

// Flux divergence of one conserved field (src/gpu_kernels.jl:165-167).  The
// split form's FP32 face fluxes are differenced in FP64 (exactly) and the
// three terms summed with FMA; the FP64 form keeps the reference's operations.
template <bool FAST>
__device__ __forceinline__ double flux_div(double fx0, double fx1, double fy0, double fy1,
                                           double fz0, double fz1, const GridDev& G) {
    if (FAST) return -__builtin_fma(fx1 - fx0, G.idx, __builtin_fma(fy1 - fy0, G.idy, (fz1 - fz0) * G.idz));
    return (-(fx1 - fx0)) * G.idx - (fy1 - fy0) * G.idy - (fz1 - fz0) * G.idz;
}

// Phi source, stage combination and floors of active cell (i, j, k) with the
// flux divergence d (src/gpu_kernels.jl:291-319, src/euler3d.jl:244-257);
// the result goes to Uout.  Returns the CFL speed bits of the new state.
template <bool FAST = false>
__device__ __forceinline__ void cell_state(const GridDev& G, int64_t c, int stage, double (&u)[5],
                                           double (&un)[5]) {
    const bool need_u = stage > 1 || G.phi != nullptr || G.src != nullptr;
    // each cell's state is read once per launch: streaming loads (the
    // stand-alone update kernel 0.515 -> 0.544 of HBM peak at 256^3)
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        un[q] = __builtin_nontemporal_load(&gmem(G.Un)[q * G.fs + c]);
        u[q] = need_u ? __builtin_nontemporal_load(&gmem(G.U)[q * G.fs + c]) : 0.0;
    }
    if (need_u && (G.lmode & 2)) apply_floor<FAST>(u, G.rho_floor, G.P_floor, G.gamma);
}

template <bool FAST = false>
__device__ __forceinline__ uint64_t finish_cell_with(const GridDev& G, int i, int j, int k, int stage,
                                                     double (&d)[5], const double (&u)[5],
                                                     const double (&un)[5], int do_cfl, double dt) {
    // sources in the order of the reference's coupled step
    // (scripts/run_sn50_fiducial.jl:772-781): point masses, self-gravity, sinks
    if (G.src && G.src->nbh > 0) bh_gravity_add(*G.src, i, j, k, u, d);
    if (G.phi) {
        const auto ph = gmem(G.phi);
        const int ip = min(i + 1, G.nx - 1), im = max(i - 1, 0);
        const int jp = min(j + 1, G.ny - 1), jm = max(j - 1, 0);
        const int kp = min(k + 1, G.nz - 1), km = max(k - 1, 0);
        const double i2x = 0.5 * G.idx, i2y = 0.5 * G.idy, i2z = 0.5 * G.idz;
        const double fxf = ip > im ? i2x : 2.0 * i2x;
        const double fyf = jp > jm ? i2y : 2.0 * i2y;
        const double fzf = kp > km ? i2z : 2.0 * i2z;
        const int64_t row = ((int64_t)k * G.ny + j) * G.nx;
        const double gx = (ph[row + ip] - ph[row + im]) * fxf;
        const double gy = (ph[((int64_t)k * G.ny + jp) * G.nx + i] -
                           ph[((int64_t)k * G.ny + jm) * G.nx + i]) * fyf;
        const double gz = (ph[((int64_t)kp * G.ny + j) * G.nx + i] -
                           ph[((int64_t)km * G.ny + j) * G.nx + i]) * fzf;
        const double r = u[0];
        const double vdotg = (u[1] * gx + u[2] * gy + u[3] * gz) / r;
        d[1] -= r * gx;
        d[2] -= r * gy;
        d[3] -= r * gz;
        d[4] -= r * vdotg;
    }
    if (G.src && G.src->nsink > 0) sink_add(*G.src, i, j, k, u, d);
    double v[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        // the split form combines with FMA (one rounding per term)
        if (FAST) {
            if (stage == 1)      v[q] = __builtin_fma(dt, d[q], un[q]);
            else if (stage == 2) v[q] = __builtin_fma(0.25 * dt, d[q], __builtin_fma(0.25, u[q], 0.75 * un[q]));
            else                 v[q] = __builtin_fma((2.0 / 3.0) * dt, d[q],
                                                      __builtin_fma(2.0 / 3.0, u[q], (1.0 / 3.0) * un[q]));
        } else {
            if (stage == 1)      v[q] = un[q] + dt * d[q];
            else if (stage == 2) v[q] = 0.75 * un[q] + 0.25 * u[q] + (0.25 * dt) * d[q];
            else                 v[q] = (1.0 / 3.0) * un[q] + (2.0 / 3.0) * u[q] + ((2.0 / 3.0) * dt) * d[q];
        }
    }
    apply_floor<FAST>(v, G.rho_floor, G.P_floor, G.gamma);
    // streaming stores: the new state is not read again by this kernel
    // (measured 1.5% faster at 256^3 than cached stores; nontemporal loads of
    // the x and y fluxes were 7% slower and are not used)
    // plane NG + k as the base, the cell's place in it as the offset
    double* const ob = G.Uout + (int64_t)(NG + k) * G.Ny * G.pitch;
    const uint32_t oo = (uint32_t)uidx(G, NG + i, NG + j, 0) * (uint32_t)sizeof(double);
#pragma unroll
    for (int q = 0; q < 5; ++q) __builtin_nontemporal_store(v[q], at_off(ob + q * G.fs, oo));
    return do_cfl ? cfl_speed_bits<FAST>(v, G) : 0;
}

template <typename T, int KZ>
__global__ __launch_bounds__(256) void k_update(const GridDev* __restrict__ grids,
                                                const Tile* __restrict__ tiles, int ntiles,
                                                int stage, int do_cfl, double* __restrict__ dU_out) {
    int t0, t1;
    chunk(ntiles, t0, t1);
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int t = t0; t < t1; ++t) {
        const Tile tl = tiles[t];
        const GridDev& G = grids[tl.g];
        const int i = tl.a + lx, j = tl.b + ly;
        uint64_t smax = 0;
        if (i < G.nx && j < G.ny) {
            const auto Fx = gmem((const T*)G.F[0]);
            const auto Fy = gmem((const T*)G.F[1]);
            const auto Fz = gmem((const T*)G.F[2]);
            const int kend = min(tl.c + KZ, G.nz);
            const double dt = stage != 0 ? G.scal[SCAL_DT] : 0.0;
            double fz0[5];
            {
                const int64_t o = ((int64_t)tl.c * G.ny + j) * G.fp[2] + i;
#pragma unroll
                for (int q = 0; q < 5; ++q) fz0[q] = (double)Fz[q * G.ffs[2] + o];
            }
            for (int k = tl.c; k < kend; ++k) {
                const int64_t ox = ((int64_t)k * G.ny + j) * G.fp[0] + i;
                const int64_t oy = ((int64_t)k * (G.ny + 1) + j) * G.fp[1] + i;
                const int64_t oz = ((int64_t)(k + 1) * G.ny + j) * G.fp[2] + i;
                const int64_t c = uidx(G, NG + i, NG + j, NG + k);
                // the cell's state is loaded with its fluxes: one round trip
                // to HBM per cell
                double u[5], un[5];
                if (stage != 0) cell_state<sizeof(T) == 4>(G, c, stage, u, un);
                double d[5], fz1[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    const double fx0 = (double)Fx[q * G.ffs[0] + ox];
                    const double fx1 = (double)Fx[q * G.ffs[0] + ox + 1];
                    const double fy0 = (double)Fy[q * G.ffs[1] + oy];
                    const double fy1 = (double)Fy[q * G.ffs[1] + oy + G.fp[1]];
                    fz1[q] = (double)Fz[q * G.ffs[2] + oz];
                    // src/gpu_kernels.jl:165-167
                    d[q] = flux_div<sizeof(T) == 4>(fx0, fx1, fy0, fy1, fz0[q], fz1[q], G);
                }
                if (stage == 0) {
#pragma unroll
                    for (int q = 0; q < 5; ++q) dU_out[q * G.fs + c] = d[q];
                } else {
                    const uint64_t b =
                        finish_cell_with<sizeof(T) == 4>(G, i, j, k, stage, d, u, un, do_cfl, dt);
                    smax = b > smax ? b : smax;
                }
#pragma unroll
                for (int q = 0; q < 5; ++q) fz0[q] = fz1[q];
            }
        }
        if (do_cfl) block_max_to(smax, (unsigned long long*)&G.scal[SCAL_SMAX]);
    }
}

// z sweep fused with the stage update.  Tile (g, i0, j0, k0 | len): 64 x 4
// threads over (i, j); each thread walks its z column from k0 for len cells with a
// six-cell register window, computes the z interface fluxes in order and
// finishes every cell as soon as both of its z faces are known, reading the
// x and y fluxes from HBM.  The z fluxes never leave registers.  The stage
// writes Uout != U, so columns can be walked in any order.
//   PF:   1 = the next window cell and the update operands are loaded before
//         the flux arithmetic (latency overlap, more registers); 2 = only the
//         next window cell is;
//   WB:   minimum waves per SIMD the register allocation must allow;
//   RING: the raw FP64 values of the last four window cells stay in a
//         per-thread LDS ring, so the update reads the cell's own state (and,
//         when Un is the stage input, its base state) from LDS instead of
//         fetching it from HBM a second time.
template <typename T, int PF, int WB, int RING>
__global__ __launch_bounds__(256, WB) void k_zupdate(const GridDev* __restrict__ grids,
                                                     const Tile* __restrict__ tiles, int ntiles,
                                                     int stage, int do_cfl) {
    constexpr bool FAST = sizeof(T) == 4;        // split form
    __shared__ double ring[RING ? 4 : 1][5][RING ? 256 : 1];
    const int rt = RING ? (int)threadIdx.x : 0;
    int t0, t1;
    chunk(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) {
        const Tile tl = tiles[t];
        const GridDev& G = grids[tl.g];
        // tile.c packs the first cell k0 (bits 0-20) and the column length (bits 21-30)
        const int k0 = tl.c & ZT_K0_MASK, klen = tl.c >> ZT_LEN_SHIFT;
        int i, j;
        col_lane(tl.a, tl.b, i, j);
        uint64_t smax = 0;
        if (i < G.nx && j < G.ny) {
            const int kend = min(k0 + klen, G.nz);
            const double dt = G.scal[SCAL_DT];
            T c[4][6], p[6];
            // interface f (between active cells f-1 and f) uses padded z = f .. f+5
            const int64_t col = uidx(G, NG + i, NG + j, 0);
            const int64_t zs = (int64_t)G.Ny * G.pitch;
            // every access below is a uniform base (field, z plane) plus the
            // lane's byte offset in the plane
            const uint32_t uo = (uint32_t)col * (uint32_t)sizeof(double);
            const uint32_t fxo = (uint32_t)(j * G.fp[0] + i) * (uint32_t)sizeof(T);
            const uint32_t fyo = (uint32_t)(j * G.fp[1] + i) * (uint32_t)sizeof(T);
            double nxt[5];
            int npz = 0;
            // the z window and the base state are read once per launch:
            // streaming loads (2.7 % on this kernel at 256^3 against cached
            // loads; the x and y fluxes, whose faces neighbouring lanes and
            // rows share, stay cached)
            auto fetch = [&](int pz) {
                npz = pz;
                // load_cell along z: a ghost plane reads the active plane its
                // index maps to; without the ring an active cell is floored here
                const int sk = (G.lmode & 1) ? bc_src(pz, G.nz, G.bc & 4) : pz;
                const double* const b = G.U + sk * zs;
#pragma unroll
                for (int q = 0; q < 5; ++q) nxt[q] = bld<double, true>(b + q * G.fs, uo);
                if (!RING && (G.lmode & 2) && pz >= NG && pz < NG + G.nz)
                    apply_floor<FAST>(nxt, G.rho_floor, G.P_floor, G.gamma);
            };
            auto place = [&](int slot) {
                if (RING) {
                    // raw value (unfloored) into the ring; slot npz & 3 is not
                    // read again before cell npz + 4 replaces it
#pragma unroll
                    for (int q = 0; q < 5; ++q) ring[npz & 3][q][rt] = nxt[q];
                    if ((G.lmode & 2) && npz >= NG && npz < NG + G.nz)
                        apply_floor<FAST>(nxt, G.rho_floor, G.P_floor, G.gamma);
                }
                const double pr = Prec<T>::pressure(nxt, G.gamma);
#pragma unroll
                for (int q = 0; q < 4; ++q) c[q][slot] = (T)nxt[q];
                p[slot] = (T)pr;
            };
            T fx[2][5], fy[2][5];
            double u[5], un[5];
            // the operands of the update of active cell k
            auto operands = [&](int k) {
                const T* const xb = (const T*)G.F[0] + (int64_t)k * G.ny * G.fp[0];
                const T* const yb = (const T*)G.F[1] + (int64_t)k * (G.ny + 1) * G.fp[1];
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    fx[0][q] = bld(xb + q * G.ffs[0], fxo);
                    fx[1][q] = bld(xb + q * G.ffs[0] + 1, fxo);
                    fy[0][q] = bld(yb + q * G.ffs[1], fyo);
                    fy[1][q] = bld(yb + q * G.ffs[1] + G.fp[1], fyo);
                }
                const int64_t cc = col + (int64_t)(NG + k) * zs;
                if (RING) {
                    const int sl = (NG + k) & 3;
                    const bool un_in_ring = G.Un == G.U;
                    const bool need_u = stage > 1 || G.phi != nullptr || G.src != nullptr;
                    const double* const nb = G.Un + (NG + k) * zs;
#pragma unroll
                    for (int q = 0; q < 5; ++q) {
                        const double r = ring[sl][q][rt];
                        u[q] = r;
                        un[q] = un_in_ring ? r : bld<double, true>(nb + q * G.fs, uo);
                    }
                    if (need_u && (G.lmode & 2)) apply_floor<FAST>(u, G.rho_floor, G.P_floor, G.gamma);
                } else {
                    cell_state<FAST>(G, cc, stage, u, un);
                }
            };
#pragma unroll
            for (int m = 0; m < 5; ++m) { fetch(k0 + m); place(m); }
            fetch(k0 + 5);
            // split form: the smoothness indicators that consecutive
            // interfaces share are carried along the column
            sp::WenoCarry wc[5];
            if constexpr (FAST) sp::interface_walk_start(c, p, wc);
            T fz0[5];                                   // the previous interface's z flux
            for (int f = k0; f <= kend; ++f) {
                place(5);                               // cell f+5
                if (PF >= 1 && f < kend) fetch(f + 6);  // next iteration's cell, in flight during the flux
                const int k = f - 1;                    // the cell finished in this iteration
                if (PF == 1 && f > k0) operands(k);
                T F[5];
                if constexpr (FAST)
                    sp::interface_flux<2, true>(c, p, G.gf32[0], G.gf32[1], G.gf32[2], F, wc);
                else
                    Prec<T>::template flux<2>(c, p, G, F);
                if (G.nzkeep > 0) {
                    // an interface the flux registers read: store it as the
                    // z-flux sweep would (adjacent tiles both compute their
                    // shared interface and store the same values)
                    bool keep = false;
                    for (int n = 0; n < G.nzkeep; ++n) keep |= G.zkeep[n] == f;
                    if (keep) {
                        const auto Fz = gmem((T*)G.F[2]);
                        const int64_t oz = ((int64_t)f * G.ny + j) * G.fp[2] + i;
#pragma unroll
                        for (int q = 0; q < 5; ++q) Fz[q * G.ffs[2] + oz] = F[q];
                    }
                }
                if (PF == 0 && f < kend) fetch(f + 6);
                if (PF != 1 && f > k0) operands(k);
                if (f > k0) {
                    double d[5];
#pragma unroll
                    for (int q = 0; q < 5; ++q)
                        d[q] = flux_div<FAST>(fx[0][q], fx[1][q], fy[0][q], fy[1][q], fz0[q], F[q], G);
                    const uint64_t b = finish_cell_with<FAST>(G, i, j, k, stage, d, u, un, do_cfl, dt);
                    smax = b > smax ? b : smax;
                }
#pragma unroll
                for (int q = 0; q < 5; ++q) fz0[q] = F[q];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[q][m] = c[q][m + 1];
                    p[m] = p[m + 1];
                }
            }
        }
        if (do_cfl) block_max_to(smax, (unsigned long long*)&G.scal[SCAL_SMAX]);
    }
}

// Stand-alone CFL maximum over active cells (cfl_dt_3d, euler3d.jl:157-169).
template <int KZ>
__global__ __launch_bounds__(256) void k_cfl(const GridDev* __restrict__ grids,
                                             const Tile* __restrict__ tiles, int ntiles) {
    int t0, t1;
    chunk(ntiles, t0, t1);
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int t = t0; t < t1; ++t) {
        const Tile tl = tiles[t];
        const GridDev& G = grids[tl.g];
        const int i = tl.a + lx, j = tl.b + ly;
        uint64_t smax = 0;
        if (i < G.nx && j < G.ny) {
            const int kend = min(tl.c + KZ, G.nz);
            for (int k = tl.c; k < kend; ++k) {
                const int64_t c = uidx(G, NG + i, NG + j, NG + k);
                double u[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) u[q] = G.U[q * G.fs + c];
                const uint64_t b = cfl_speed_bits(u, G);
                smax = b > smax ? b : smax;
            }
        }
        block_max_to(smax, (unsigned long long*)&G.scal[SCAL_SMAX]);
    }
}

// dt = cfl / smax (Inf when smax is not > 0), optionally clipped to t_end - t;
// the maximum is reset for the next accumulation.
__global__ void k_dt(const GridDev* __restrict__ grids, int ngrids) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrids) return;
    double* s = grids[g].scal;
    const uint64_t bits = *(const uint64_t*)&s[SCAL_SMAX];
    const double smax = __longlong_as_double((long long)bits);
    double dt = smax > 0.0 ? s[SCAL_CFL] / smax : __builtin_inf();
    if (s[SCAL_TEND] > 0.0) dt = jl_min(dt, s[SCAL_TEND] - s[SCAL_T]);
    s[SCAL_DT] = dt;
    *(uint64_t*)&s[SCAL_SMAX] = 0;
}

__global__ void k_advance_t(const GridDev* __restrict__ grids, int ngrids) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrids) return;
    double* s = grids[g].scal;
    s[SCAL_T] = s[SCAL_T] + s[SCAL_DT];
}

__global__ void k_set_scal(double* scal, int idx, double v) { scal[idx] = v; }

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        p[e] = v;
}

// ---------------------------------------------------------------------------
// Layout conversion between the reference's (5, Nx, Ny, Nz) column-major
// array and the device layout.  dst/src select the device buffer (U, Un or
// any buffer with U's layout).

__global__ void k_aos_to_soa(GridDev G, const double* __restrict__ aos, double* __restrict__ dst) {
    const int Nx = G.nx + 2 * NG;
    const int64_t n = (int64_t)Nx * G.Ny * G.Nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Nx);
        const int64_t r = e / Nx;
        const int j = (int)(r % G.Ny), k = (int)(r / G.Ny);
        const int64_t c = uidx(G, i, j, k);
#pragma unroll
        for (int q = 0; q < 5; ++q) dst[q * G.fs + c] = aos[5 * e + q];
    }
}

__global__ void k_soa_to_aos(GridDev G, const double* __restrict__ src, double* __restrict__ aos) {
    const int Nx = G.nx + 2 * NG;
    const int64_t n = (int64_t)Nx * G.Ny * G.Nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Nx);
        const int64_t r = e / Nx;
        const int j = (int)(r % G.Ny), k = (int)(r / G.Ny);
        const int64_t c = uidx(G, i, j, k);
#pragma unroll
        for (int q = 0; q < 5; ++q) aos[5 * e + q] = src[q * G.fs + c];
    }
}

// Fluxes to the reference's F? arrays (zero outside the active-range
// interfaces, as after euler3d.jl:197-202).  Shapes: Fx (5, Nx+1, Ny, Nz),
// Fy (5, Nx, Ny+1, Nz), Fz (5, Nx, Ny, Nz+1).
template <typename T>
__global__ void k_flux_export(GridDev G, int dir, double* __restrict__ out) {
    const int Nx = G.nx + 2 * NG;
    const int Ex = Nx + (dir == 0), Ey = G.Ny + (dir == 1), Ez = G.Nz + (dir == 2);
    const int64_t n = (int64_t)Ex * Ey * Ez;
    const T* F = (const T*)G.F[dir];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(e % Ex);
        const int64_t r = e / Ex;
        const int b = (int)(r % Ey), c = (int)(r / Ey);
        // reference 0-based index of interface f is NG - 1 + f along dir
        int f = -1, ci = -1, cj = -1;
        bool in;
        if (dir == 0) { f = a - (NG - 1); ci = b - NG; cj = c - NG;
                        in = f >= 0 && f <= G.nx && ci >= 0 && ci < G.ny && cj >= 0 && cj < G.nz; }
        else if (dir == 1) { f = b - (NG - 1); ci = a - NG; cj = c - NG;
                        in = f >= 0 && f <= G.ny && ci >= 0 && ci < G.nx && cj >= 0 && cj < G.nz; }
        else { f = c - (NG - 1); ci = a - NG; cj = b - NG;
               in = f >= 0 && f <= G.nz && ci >= 0 && ci < G.nx && cj >= 0 && cj < G.ny; }
        int64_t o = 0;
        if (in) {
            if (dir == 0)      o = ((int64_t)cj * G.ny + ci) * G.fp[0] + f;
            else if (dir == 1) o = ((int64_t)cj * (G.ny + 1) + f) * G.fp[1] + ci;
            else               o = ((int64_t)f * G.ny + cj) * G.fp[2] + ci;
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) out[5 * e + q] = in ? (double)F[q * G.ffs[dir] + o] : 0.0;
    }
}

// Phi source on a dU buffer with U's layout (_apply_sg_force_level!,
// fmr_gravity.jl:146-156 -> gpu_kernels.jl:291-319).
__global__ void k_sg_force(GridDev G, double* __restrict__ dU) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % G.nx);
        const int64_t r = e / G.nx;
        const int j = (int)(r % G.ny), k = (int)(r / G.ny);
        const int64_t c = uidx(G, NG + i, NG + j, NG + k);
        const double* ph = G.phi;
        const int ip = min(i + 1, G.nx - 1), im = max(i - 1, 0);
        const int jp = min(j + 1, G.ny - 1), jm = max(j - 1, 0);
        const int kp = min(k + 1, G.nz - 1), km = max(k - 1, 0);
        const double i2x = 0.5 * G.idx, i2y = 0.5 * G.idy, i2z = 0.5 * G.idz;
        const double fxf = ip > im ? i2x : 2.0 * i2x;
        const double fyf = jp > jm ? i2y : 2.0 * i2y;
        const double fzf = kp > km ? i2z : 2.0 * i2z;
        const double gx = (ph[((int64_t)k * G.ny + j) * G.nx + ip] - ph[((int64_t)k * G.ny + j) * G.nx + im]) * fxf;
        const double gy = (ph[((int64_t)k * G.ny + jp) * G.nx + i] - ph[((int64_t)k * G.ny + jm) * G.nx + i]) * fyf;
        const double gz = (ph[((int64_t)kp * G.ny + j) * G.nx + i] - ph[((int64_t)km * G.ny + j) * G.nx + i]) * fzf;
        const double rr = G.U[c], mx = G.U[G.fs + c], my = G.U[2 * G.fs + c], mz = G.U[3 * G.fs + c];
        const double vdotg = (mx * gx + my * gy + mz * gz) / rr;
        dU[G.fs + c] -= rr * gx;
        dU[2 * G.fs + c] -= rr * gy;
        dU[3 * G.fs + c] -= rr * gz;
        dU[4 * G.fs + c] -= rr * vdotg;
    }
}

// Point-mass gravity (which = 0) or sink (which = 1) source on a dU buffer
// with U's layout, from the grid's state (add_bh_gravity_source!,
// src/gravity_bh.jl:58-78; add_sink_sources!, src/sinks.jl:32-58).
__global__ void k_src(GridDev G, double* __restrict__ dU, int which) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % G.nx);
        const int64_t r = e / G.nx;
        const int j = (int)(r % G.ny), k = (int)(r / G.ny);
        const int64_t c = uidx(G, NG + i, NG + j, NG + k);
        double u[5], d[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            u[q] = G.U[q * G.fs + c];
            d[q] = dU[q * G.fs + c];
        }
        if (which == 0) bh_gravity_add(*G.src, i, j, k, u, d);
        else            sink_add(*G.src, i, j, k, u, d);
#pragma unroll
        for (int q = 0; q < 5; ++q) dU[q * G.fs + c] = d[q];
    }
}

// Whole-array RK3 combine with a caller-supplied dU in U's layout
// (rk3.jl:24,28,32 / euler3d.jl:244,250,256), over every padded cell.
__global__ void k_combine_all(GridDev G, int stage, const double* __restrict__ dU, double dt) {
    const int Nx = G.nx + 2 * NG;
    const int64_t n = (int64_t)Nx * G.Ny * G.Nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Nx);
        const int64_t r = e / Nx;
        const int j = (int)(r % G.Ny), k = (int)(r / G.Ny);
        const int64_t c = uidx(G, i, j, k);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const double un = G.Un[q * G.fs + c], u = G.U[q * G.fs + c], d = dU[q * G.fs + c];
            double v;
            if (stage == 1)      v = un + dt * d;
            else if (stage == 2) v = 0.75 * un + 0.25 * u + (0.25 * dt) * d;
            else                 v = (1.0 / 3.0) * un + (2.0 / 3.0) * u + ((2.0 / 3.0) * dt) * d;
            G.U[q * G.fs + c] = v;
        }
    }
}

// Copy every padded cell of src into dst (both with U's layout).
__global__ void k_copy_state(GridDev G, const double* __restrict__ src, double* __restrict__ dst) {
    const int64_t n = 5 * G.fs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

// Conservation totals: per-block partial sums over active cells, then a
// fixed-order final sum (deterministic for a fixed launch shape).
__global__ void k_totals_partial(GridDev G, double* __restrict__ part) {
    __shared__ double red[5][256];
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % G.nx);
        const int64_t r = e / G.nx;
        const int j = (int)(r % G.ny), k = (int)(r / G.ny);
        const int64_t c = uidx(G, NG + i, NG + j, NG + k);
#pragma unroll
        for (int q = 0; q < 5; ++q) acc[q] += G.U[q * G.fs + c];
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 5; ++q) part[5 * blockIdx.x + q] = red[q][0];
}

__global__ void k_totals_final(const double* __restrict__ part, int nblocks, double dV,
                               double* __restrict__ out) {
    const int q = threadIdx.x;
    if (q >= 5) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[5 * b + q];
    out[q] = s * dV;
}

// ---------------------------------------------------------------------------
// FMR transfer operators (src/fmr3d.jl).  Coarse index arithmetic follows
// _coarse_stencil (fmr3d.jl:165-170) in FP64 with ties-to-even rounding.

__device__ __forceinline__ void coarse_stencil(int ig_f1, int c_lo, int ratio, int& ic0, int& sx) {
    // ig_f1 is the reference's 1-based fine index
    const double xi = ((double)c_lo - 0.5) + ((double)(ig_f1 - NG) - 0.5) / (double)ratio;
    const double r = rint(xi);
    ic0 = (int)r;
    int s = (int)floor(((xi - r) + 0.5) * (double)ratio);
    sx = s < 0 ? 0 : (s > ratio - 1 ? ratio - 1 : s);
}

// Prolongation of fine cells from the parent state interpolated in time,
// Uc(alpha) = (1 - alpha) Ucn + alpha Ucnp1 (fmr3d.jl:551), evaluated per
// coarse value on the fly.  all_cells = 1 fills every fine cell
// (_prolong_all!), else only ghost cells in the reference's slab partition.
// The values go to each of the dst.n buffers in the child's layout, or to
// the child's state when dst.n is 0.
__global__ void k_prolong(const GridDev* __restrict__ grids, const FmrDev* __restrict__ fmr,
                          const double* __restrict__ Uca, const double* __restrict__ Ucat,
                          int all_cells, ProlongDst dst) {
    const FmrDev& M = *fmr;
    const GridDev& C = grids[M.parent];
    const GridDev& Fg = grids[M.child];
    const int Nfx = Fg.nx + 2 * NG;
    const int64_t total = all_cells ? (int64_t)Nfx * Fg.Ny * Fg.Nz : ghost_count(Fg);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        // x-slab ghost cells read the transposed copy when there is one: their
        // lanes then run along y, where consecutive cells are adjacent in it
        const bool xt = !all_cells && Ucat && e < 2 * NG * (int64_t)Fg.Ny * Fg.Nz;
        if (all_cells) {
            i = (int)(e % Nfx);
            const int64_t r = e / Nfx;
            j = (int)(r % Fg.Ny); k = (int)(r / Fg.Ny);
        } else if (xt) {
            j = (int)(e % Fg.Ny);
            const int64_t r = e / Fg.Ny;
            k = (int)(r % Fg.Nz);
            const int ig = (int)(r / Fg.Nz);
            i = ig < NG ? ig : Fg.nx + ig;
        } else {
            ghost_decode(Fg, e, i, j, k);
        }
        int ic0, jc0, kc0, si, sj, sk;
        coarse_stencil(i + 1, M.lo[0], M.ratio, ic0, si);
        coarse_stencil(j + 1, M.lo[1], M.ratio, jc0, sj);
        coarse_stencil(k + 1, M.lo[2], M.ratio, kc0, sk);
        const int64_t d = uidx(Fg, i, j, k);
        if (xt) {
            // the same terms in the same order, from the transposed box
            const int bx = M.phi[0] - M.plo[0] + 1, by = M.phi[1] - M.plo[1] + 1;
            const int bz = M.phi[2] - M.plo[2] + 1;
            const int64_t bt = (int64_t)bx * by * bz;
            double val[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            for (int di = -2; di <= 2; ++di) {
                const int ig = NG + ic0 + di - 1 - M.plo[0];
                const double wi = M.w[si * 5 + di + 2];
                for (int dj = -2; dj <= 2; ++dj) {
                    const int jg = NG + jc0 + dj - 1 - M.plo[1];
                    const double wij = wi * M.w[sj * 5 + dj + 2];
                    const int64_t ob = ((int64_t)ig * bz + (NG + kc0 - 3 - M.plo[2])) * by + jg;
#pragma unroll
                    for (int dk = 0; dk < 5; ++dk) {
                        const int64_t o = ob + (int64_t)dk * by;
                        const double w = wij * M.w[sk * 5 + dk];
#pragma unroll
                        for (int q = 0; q < 5; ++q) val[q] += w * Ucat[q * bt + o];
                    }
                }
            }
            if (dst.n == 0) {
#pragma unroll
                for (int q = 0; q < 5; ++q) Fg.U[q * Fg.fs + d] = val[q];
            }
            for (int b = 0; b < dst.n; ++b) {
#pragma unroll
                for (int q = 0; q < 5; ++q) dst.p[b][q * Fg.fs + d] = val[q];
            }
            continue;
        }
        const int64_t zs = (int64_t)C.Ny * C.pitch;
        // the five fields accumulate side by side; each field's terms are
        // summed in the reference's (di, dj, dk) order
        double val[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int di = -2; di <= 2; ++di) {
            const int ig = NG + ic0 + di - 1;
            const double wi = M.w[si * 5 + di + 2];
            for (int dj = -2; dj <= 2; ++dj) {
                const int jg = NG + jc0 + dj - 1;
                const double wij = wi * M.w[sj * 5 + dj + 2];
                const int64_t ob = uidx(C, ig, jg, NG + kc0 - 3);
#pragma unroll
                for (int dk = 0; dk < 5; ++dk) {
                    const int64_t o = ob + dk * zs;
                    const double w = wij * M.w[sk * 5 + dk];
#pragma unroll
                    for (int q = 0; q < 5; ++q) val[q] += w * Uca[q * C.fs + o];
                }
            }
        }
        if (dst.n == 0) {
#pragma unroll
            for (int q = 0; q < 5; ++q) Fg.U[q * Fg.fs + d] = val[q];
        }
        for (int b = 0; b < dst.n; ++b) {
#pragma unroll
            for (int q = 0; q < 5; ++q) dst.p[b][q * Fg.fs + d] = val[q];
        }
    }
}

// U_c_interp = (1 - alpha) * U_c_n + alpha * U_c_np1 (fmr3d.jl:547-551) over
// the parent cells the child's prolongation stencils read, once per fine
// sub-cycle (alpha is fixed for its three stages).
__global__ void k_time_interp(const GridDev* __restrict__ grids, const FmrDev* __restrict__ fmr,
                              double alpha) {
    const FmrDev& M = *fmr;
    const GridDev& C = grids[M.parent];
    const int nx = M.phi[0] - M.plo[0] + 1, ny = M.phi[1] - M.plo[1] + 1;
    const int64_t total = (int64_t)nx * ny * (M.phi[2] - M.plo[2] + 1);
    const double oma = 1.0 - alpha;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = M.plo[0] + (int)(e % nx);
        const int64_t r = e / nx;
        const int j = M.plo[1] + (int)(r % ny), k = M.plo[2] + (int)(r / ny);
        const int64_t o = uidx(C, i, j, k);
        const int nz = M.phi[2] - M.plo[2] + 1;
        const int64_t ot = ((int64_t)(i - M.plo[0]) * nz + (k - M.plo[2])) * ny + (j - M.plo[1]);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const double v = oma * M.Ucn[q * C.fs + o] + alpha * M.Ucnp1[q * C.fs + o];
            M.Uci[q * C.fs + o] = v;
            M.Ucit[q * total + ot] = v;
        }
    }
}

// _restrict! (fmr3d.jl:266-285)
// _restrict! (fields 0..nq-1; nq = 5) and restrict_density_to_coarse!
// (nq = 1, src/fmr_gravity.jl:40-60): the same sum in the same order.
__global__ void k_restrict(const GridDev* __restrict__ grids, const FmrDev* __restrict__ fmr, int nq) {
    const FmrDev& M = *fmr;
    const GridDev& C = grids[M.parent];
    const GridDev& Fg = grids[M.child];
    const int r = M.ratio;
    const int nic = M.hi[0] - M.lo[0] + 1, njc = M.hi[1] - M.lo[1] + 1, nkc = M.hi[2] - M.lo[2] + 1;
    const int64_t total = (int64_t)nic * njc * nkc;
    const double inv_r3 = 1.0 / (double)(r * r * r);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(e % nic);
        const int64_t rr = e / nic;
        const int n = (int)(rr % njc), p = (int)(rr / njc);
        const int64_t dc = uidx(C, NG + M.lo[0] + m - 1, NG + M.lo[1] + n - 1, NG + M.lo[2] + p - 1);
        for (int q = 0; q < nq; ++q) {
            double val = 0.0;
            for (int sx = 1; sx <= r; ++sx)
                for (int sy = 1; sy <= r; ++sy)
                    for (int sz = 1; sz <= r; ++sz)
                        val += Fg.U[q * Fg.fs + uidx(Fg, NG + r * m + sx - 1, NG + r * n + sy - 1,
                                                     NG + r * p + sz - 1)];
            C.U[q * C.fs + dc] = val * inv_r3;
        }
    }
}

// _prolong_scalar_trilinear! (src/fmr_gravity.jl:70-106): phi of the fine
// active cells (x fastest) from the coarse active-cell phi, the fine box
// starting at coarse active cell (lo_x, lo_y, lo_z) (1-based), in the
// reference's operation order.
__global__ void k_prolong_scalar(const double* __restrict__ phc, int ncx, int ncy, int ncz,
                                 double* __restrict__ phf, int nfx, int nfy, int nfz, int lox,
                                 int loy, int loz, int r) {
    const int64_t total = (int64_t)nfx * nfy * nfz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % nfx);
        const int64_t rr = e / nfx;
        const int j = (int)(rr % nfy), k = (int)(rr / nfy);
        // fine cell centre in coarse active-cell coordinates (1-based)
        const double xc = ((double)(lox - 1) + ((double)(i + 1) - 0.5) / (double)r) + 0.5;
        const double yc = ((double)(loy - 1) + ((double)(j + 1) - 0.5) / (double)r) + 0.5;
        const double zc = ((double)(loz - 1) + ((double)(k + 1) - 0.5) / (double)r) + 0.5;
        const int ic = min(max((int)floor(xc), 1), ncx - 1);
        const int jc = min(max((int)floor(yc), 1), ncy - 1);
        const int kc = min(max((int)floor(zc), 1), ncz - 1);
        const double fx = xc - ic, fy = yc - jc, fz = zc - kc;
        const double w000 = (1 - fx) * (1 - fy) * (1 - fz);
        const double w100 = fx * (1 - fy) * (1 - fz);
        const double w010 = (1 - fx) * fy * (1 - fz);
        const double w110 = fx * fy * (1 - fz);
        const double w001 = (1 - fx) * (1 - fy) * fz;
        const double w101 = fx * (1 - fy) * fz;
        const double w011 = (1 - fx) * fy * fz;
        const double w111 = fx * fy * fz;
        auto P = [&](int a, int b, int c) {
            return phc[((int64_t)(kc - 1 + c) * ncy + (jc - 1 + b)) * ncx + (ic - 1 + a)];
        };
        phf[e] = w000 * P(0, 0, 0) + w100 * P(1, 0, 0) + w010 * P(0, 1, 0) + w110 * P(1, 1, 0) +
                 w001 * P(0, 0, 1) + w101 * P(1, 0, 1) + w011 * P(0, 1, 1) + w111 * P(1, 1, 1);
    }
}

// a .*= s
__global__ void k_scale(double* __restrict__ a, int64_t n, double s) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        a[e] = a[e] * s;
}

// Flux registers: face f in 0..5 = x_lo, x_hi, y_lo, y_hi, z_lo, z_hi, each
// (5, n1, n2) column-major as FluxReg3D (fmr3d.jl:81-96).
__device__ __forceinline__ void face_dims(const FmrDev& M, int face, int& n1, int& n2) {
    const int nic = M.hi[0] - M.lo[0] + 1, njc = M.hi[1] - M.lo[1] + 1, nkc = M.hi[2] - M.lo[2] + 1;
    if (face < 2)      { n1 = njc; n2 = nkc; }
    else if (face < 4) { n1 = nic; n2 = nkc; }
    else               { n1 = nic; n2 = njc; }
}

__global__ void k_fr_zero(const FmrDev* __restrict__ fmr) {
    const FmrDev& M = *fmr;
    const int64_t n = M.froff[6];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        M.frc[e] = 0.0;
        M.frf[e] = 0.0;
    }
}

template <typename T>
__device__ __forceinline__ double flux_at(const GridDev& G, int dir, int q, int f, int a, int b) {
    // dir 0: (f, j=a, k=b); dir 1: (i=a, f, k=b); dir 2: (i=a, j=b, f)
    const T* F = (const T*)G.F[dir];
    int64_t o;
    if (dir == 0)      o = ((int64_t)b * G.ny + a) * G.fp[0] + f;
    else if (dir == 1) o = ((int64_t)b * (G.ny + 1) + f) * G.fp[1] + a;
    else               o = ((int64_t)f * G.ny + b) * G.fp[2] + a;
    return (double)F[q * G.ffs[dir] + o];
}

// _accumulate_fr_coarse! (fmr3d.jl:292-326); fac = weight * dt of the parent.
template <typename T>
__global__ void k_fr_coarse(const GridDev* __restrict__ grids, const FmrDev* __restrict__ fmr,
                            double weight) {
    const int q = blockIdx.y;                   // one field per grid row
    const FmrDev& M = *fmr;
    const GridDev& C = grids[M.parent];
    const double fac = weight * C.scal[SCAL_DT];
    const int64_t n = M.froff[6] / 5;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        int face = 0;
        while (face < 5 && e >= M.froff[face + 1] / 5) ++face;
        const int64_t loc = e - M.froff[face] / 5;
        int n1, n2;
        face_dims(M, face, n1, n2);
        const int a = (int)(loc % n1), b = (int)(loc / n1);
        const int hi = face & 1;
        // reference interface NG + ci_lo - 1 (lo) or NG + ci_hi (hi) is
        // device interface f = ci_lo - 1 or ci_hi (active-cell faces)
        double F;
        if (face < 2)      F = flux_at<T>(C, 0, q, hi ? M.hi[0] : M.lo[0] - 1, M.lo[1] - 1 + a, M.lo[2] - 1 + b);
        else if (face < 4) F = flux_at<T>(C, 1, q, hi ? M.hi[1] : M.lo[1] - 1, M.lo[0] - 1 + a, M.lo[2] - 1 + b);
        else               F = flux_at<T>(C, 2, q, hi ? M.hi[2] : M.lo[2] - 1, M.lo[0] - 1 + a, M.lo[1] - 1 + b);
        double* R = &M.frc[M.froff[face] + q + 5 * loc];
        *R = *R + fac * F;
    }
}

// _accumulate_fr_fine! (fmr3d.jl:331-369): each register element sums its
// r^2 fine faces in the reference's loop order; fac = weight*dt_f/r^2.
template <typename T>
__global__ void k_fr_fine(const GridDev* __restrict__ grids, const FmrDev* __restrict__ fmr,
                          double weight) {
    const int q = blockIdx.y;                   // one field per grid row
    const FmrDev& M = *fmr;
    const GridDev& Fg = grids[M.child];
    const int r = M.ratio;
    const double fac = (weight * Fg.scal[SCAL_DT]) / (double)(r * r);
    const int64_t n = M.froff[6] / 5;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        int face = 0;
        while (face < 5 && e >= M.froff[face + 1] / 5) ++face;
        const int64_t loc = e - M.froff[face] / 5;
        int n1, n2;
        face_dims(M, face, n1, n2);
        const int a = (int)(loc % n1), b = (int)(loc / n1);
        const int hi = face & 1;
        double* R = &M.frf[M.froff[face] + q + 5 * loc];
        double acc = *R;
        // outer loop over the second register index (kf or jf), inner over
        // the first (jf or ia), as the reference
        for (int ob = 0; ob < r; ++ob)
            for (int oa = 0; oa < r; ++oa) {
                const int fa = a * r + oa, fb = b * r + ob;   // 0-based fine active
                double F;
                if (face < 2)      F = flux_at<T>(Fg, 0, q, hi ? Fg.nx : 0, fa, fb);
                else if (face < 4) F = flux_at<T>(Fg, 1, q, hi ? Fg.ny : 0, fa, fb);
                else               F = flux_at<T>(Fg, 2, q, hi ? Fg.nz : 0, fa, fb);
                acc = acc + fac * F;
            }
        *R = acc;
    }
}

// _bc_correct! (fmr3d.jl:379-422): every corrected coarse cell receives one
// face's correction, so the faces are applied in parallel.
__global__ void k_bc_correct(const GridDev* __restrict__ grids, const FmrDev* __restrict__ fmr) {
    const FmrDev& M = *fmr;
    const GridDev& C = grids[M.parent];
    const int64_t n = M.froff[6] / 5;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        int face = 0;
        while (face < 5 && e >= M.froff[face + 1] / 5) ++face;
        const int64_t loc = e - M.froff[face] / 5;
        int n1, n2;
        face_dims(M, face, n1, n2);
        const int a = (int)(loc % n1), b = (int)(loc / n1);
        const int hi = face & 1;
        int ci, cj, ck;   // 1-based coarse active
        double inv;
        if (face < 2)      { ci = hi ? M.hi[0] + 1 : M.lo[0] - 1; cj = M.lo[1] + a; ck = M.lo[2] + b; inv = C.idx; }
        else if (face < 4) { ci = M.lo[0] + a; cj = hi ? M.hi[1] + 1 : M.lo[1] - 1; ck = M.lo[2] + b; inv = C.idy; }
        else               { ci = M.lo[0] + a; cj = M.lo[1] + b; ck = hi ? M.hi[2] + 1 : M.lo[2] - 1; inv = C.idz; }
        const int64_t c = uidx(C, NG + ci - 1, NG + cj - 1, NG + ck - 1);
        for (int q = 0; q < 5; ++q) {
            const double rc = M.frc[M.froff[face] + q + 5 * loc];
            const double rf = M.frf[M.froff[face] + q + 5 * loc];
            const double delta = hi ? (rf - rc) * inv : (rc - rf) * inv;
            C.U[q * C.fs + c] += delta;
        }
    }
}

// dt of a nested level: dt_child = dt_parent / ratio (fmr3d.jl:454).
__global__ void k_child_dt(const GridDev* __restrict__ grids, int parent, int child, int ratio) {
    grids[child].scal[SCAL_DT] = grids[parent].scal[SCAL_DT] / (double)ratio;
}

// dt_c = min(dt_c, r * dt_f) over a hierarchy (cfl_dt_fmr3d, fmr3d.jl:616-622):
// level dts are computed first by k_dt on every grid, then folded into the root.
__global__ void k_fold_dt(const GridDev* __restrict__ grids, int root, int child, double factor) {
    double* s = grids[root].scal;
    s[SCAL_DT] = jl_min(s[SCAL_DT], factor * grids[child].scal[SCAL_DT]);
}

// ---------------------------------------------------------------------------
// Launch helpers

static int resident_blocks(const void* fn, int threads, size_t lds) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds) != hipSuccess || per < 1)
        per = 1;
    return per;
}

static int occupancy_grid(const void* fn, int threads, size_t lds, int ntiles, int cus) {
    int64_t g = (int64_t)resident_blocks(fn, threads, lds) * cus;
    if (g > ntiles) g = ntiles;
    return (int)(g < 1 ? 1 : g);
}

template <int DIR, typename T, int W, int H>
static hipError_t launch_flux_t(const LaunchCtx& L, const GridDev* grids, const Tile* tiles,
                                int ntiles, int max_wg) {
    if (ntiles <= 0) return hipSuccess;
    const void* fn = (const void*)&k_flux<DIR, T, W, H>;
    int g = occupancy_grid(fn, W * H, 0, ntiles, L.cus);
    if (max_wg > 0 && g > max_wg) g = max_wg;
    hipLaunchKernelGGL((k_flux<DIR, T, W, H>), dim3(g), dim3(W * H), 0, L.stream, grids, tiles, ntiles);
    return hipGetLastError();
}

template <int DIR, typename T>
static hipError_t launch_fwalk_t(const LaunchCtx& L, const GridDev* grids, const Tile* tiles,
                                 int ntiles, int max_wg) {
    if (ntiles <= 0) return hipSuccess;
    const void* fn = (const void*)&k_fwalk<DIR, T>;
    int g = occupancy_grid(fn, 256, 0, ntiles, L.cus);
    if (max_wg > 0 && g > max_wg) g = max_wg;
    hipLaunchKernelGGL((k_fwalk<DIR, T>), dim3(g), dim3(256), 0, L.stream, grids, tiles, ntiles);
    return hipGetLastError();
}

template <typename T, int KMAX, bool OUT>
static hipError_t launch_fxwalk_t(const LaunchCtx& L, const GridDev* grids, const Tile* tiles,
                                  int ntiles, int max_wg) {
    if (ntiles <= 0) return hipSuccess;
    const void* fn = (const void*)&k_fxwalk<T, KMAX, OUT>;
    int g = occupancy_grid(fn, 256, 0, ntiles, L.cus);
    if (max_wg > 0 && g > max_wg) g = max_wg;
    hipLaunchKernelGGL((k_fxwalk<T, KMAX, OUT>), dim3(g), dim3(256), 0, L.stream, grids, tiles, ntiles);
    return hipGetLastError();
}

template <int DIR, typename T>
static hipError_t launch_flux_shape(const LaunchCtx& L, int shape, const GridDev* grids,
                                    const Tile* tiles, int ntiles, int max_wg) {
    if (DIR != 0 && shape >= FLUX_WALK_SHAPE0)
        return launch_fwalk_t<DIR == 0 ? 1 : DIR, T>(L, grids, tiles, ntiles, max_wg);
    if (DIR == 0 && shape >= FLUX_XWALK_SHAPE0)
        switch (shape - FLUX_XWALK_SHAPE0) {
            case 0: return launch_fxwalk_t<T, 2, true>(L, grids, tiles, ntiles, max_wg);
            case 1: return launch_fxwalk_t<T, 4, true>(L, grids, tiles, ntiles, max_wg);
            default: return launch_fxwalk_t<T, 4, false>(L, grids, tiles, ntiles, max_wg);
        }
    // shape ids: x sweep 0: 64x4, 1: 32x8, 2: 16x16;  y/z sweep 0: 64x4, 1: 64x8, 2: 32x8, 3: 64x2,
    // 4: 32x16
    if (DIR == 0) {
        switch (shape) {
            case 1: return launch_flux_t<DIR, T, 32, 8>(L, grids, tiles, ntiles, max_wg);
            case 2: return launch_flux_t<DIR, T, 16, 16>(L, grids, tiles, ntiles, max_wg);
            default: return launch_flux_t<DIR, T, 64, 4>(L, grids, tiles, ntiles, max_wg);
        }
    } else {
        switch (shape) {
            case 1: return launch_flux_t<DIR, T, 64, 8>(L, grids, tiles, ntiles, max_wg);
            case 2: return launch_flux_t<DIR, T, 32, 8>(L, grids, tiles, ntiles, max_wg);
            case 3: return launch_flux_t<DIR, T, 64, 2>(L, grids, tiles, ntiles, max_wg);
            case 4: return launch_flux_t<DIR, T, 32, 16>(L, grids, tiles, ntiles, max_wg);
            default: return launch_flux_t<DIR, T, 64, 4>(L, grids, tiles, ntiles, max_wg);
        }
    }
}

// W x H of a tiled shape; for the walk shapes W = 64 lanes and H = 4 rows or
// columns, with the run length flux_walk_len(dir, shape).
void flux_shape_dims(int dir, int shape, int& W, int& H) {
    static const int xs[3][2] = {{64, 4}, {32, 8}, {16, 16}};
    static const int ys[5][2] = {{64, 4}, {64, 8}, {32, 8}, {64, 2}, {32, 16}};
    if (flux_walk_len(dir, shape) > 0) { W = 64; H = 4; return; }
    if (dir == 0) { if (shape < 0 || shape > 2) shape = 0; W = xs[shape][0]; H = xs[shape][1]; }
    else          { if (shape < 0 || shape > 4) shape = 0; W = ys[shape][0]; H = ys[shape][1]; }
}

// Run length of a walk shape (interfaces per thread; for x the most one lane
// takes), 0 for a tiled shape.
int flux_walk_len(int dir, int shape) {
    if (shape < 0 || shape >= flux_shape_count(dir)) return 0;
    if (dir == 0) {
        static const int lens[3] = {2, 4, 4};
        return shape >= FLUX_XWALK_SHAPE0 ? lens[shape - FLUX_XWALK_SHAPE0] : 0;
    }
    static const int lens[3] = {16, 32, 64};
    return shape >= FLUX_WALK_SHAPE0 ? lens[shape - FLUX_WALK_SHAPE0] : 0;
}

int flux_slots(int dir, int prec, int shape, int cus) {
    // only the walks' tile lists depend on it
    if (dir == 0 || flux_walk_len(dir, shape) <= 0) return cus;
    const void* fn = prec == 0 ? (dir == 1 ? (const void*)&k_fwalk<1, double> : (const void*)&k_fwalk<2, double>)
                               : (dir == 1 ? (const void*)&k_fwalk<1, float> : (const void*)&k_fwalk<2, float>);
    return resident_blocks(fn, 256, 0) * cus;
}

int flux_shape_count(int dir) { return dir == 0 ? FLUX_XWALK_SHAPE0 + 3 : FLUX_WALK_SHAPE0 + 3; }

hipError_t launch_flux(const LaunchCtx& L, int dir, int prec, int shape, const GridDev* grids,
                       const Tile* tiles, int ntiles, int max_wg) {
    if (prec == 0) {
        if (dir == 0) return launch_flux_shape<0, double>(L, shape, grids, tiles, ntiles, max_wg);
        if (dir == 1) return launch_flux_shape<1, double>(L, shape, grids, tiles, ntiles, max_wg);
        return launch_flux_shape<2, double>(L, shape, grids, tiles, ntiles, max_wg);
    }
    if (dir == 0) return launch_flux_shape<0, float>(L, shape, grids, tiles, ntiles, max_wg);
    if (dir == 1) return launch_flux_shape<1, float>(L, shape, grids, tiles, ntiles, max_wg);
    return launch_flux_shape<2, float>(L, shape, grids, tiles, ntiles, max_wg);
}

hipError_t launch_update(const LaunchCtx& L, int prec, const GridDev* grids, const Tile* tiles,
                         int ntiles, int stage, int do_cfl, double* dU_out) {
    if (ntiles <= 0) return hipSuccess;
    if (prec == 0) {
        const void* fn = (const void*)&k_update<double, UPDATE_KZ>;
        const int g = occupancy_grid(fn, 256, 0, ntiles, L.cus);
        hipLaunchKernelGGL((k_update<double, UPDATE_KZ>), dim3(g), dim3(256), 0, L.stream, grids,
                           tiles, ntiles, stage, do_cfl, dU_out);
    } else {
        const void* fn = (const void*)&k_update<float, UPDATE_KZ>;
        const int g = occupancy_grid(fn, 256, 0, ntiles, L.cus);
        hipLaunchKernelGGL((k_update<float, UPDATE_KZ>), dim3(g), dim3(256), 0, L.stream, grids,
                           tiles, ntiles, stage, do_cfl, dU_out);
    }
    return hipGetLastError();
}

static int flat_blocks(int64_t n, int cus);

template <typename T, int PF, int WB, int RING>
static hipError_t launch_zupdate_t(const LaunchCtx& L, const GridDev* grids, const Tile* tiles,
                                   int ntiles, int stage, int do_cfl) {
    const void* fn = (const void*)&k_zupdate<T, PF, WB, RING>;
    const int g = occupancy_grid(fn, 256, 0, ntiles, L.cus);
    hipLaunchKernelGGL((k_zupdate<T, PF, WB, RING>), dim3(g), dim3(256), 0, L.stream, grids,
                       tiles, ntiles, stage, do_cfl);
    return hipGetLastError();
}

template <typename T, int PF, int WB, int RING>
static int zupdate_slots_t(int cus) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)&k_zupdate<T, PF, WB, RING>,
                                                     256, 0) != hipSuccess || per < 1)
        per = 1;
    return per * cus;
}

// Shapes: 3 * variant + segmentation.  Variant 0 plain, 1 operands loaded one
// cell ahead (FP64) or variant 3 without a register cap (split: 2 waves per
// SIMD), 2 LDS ring of the window's raw cells, 3 ring plus an early load of
// the next window cell (split: both capped at 168 VGPRs, 3 waves per SIMD).  Segmentation 0: 16 cells per column tile, 1: 32,
// 2: one length per grid chosen so the tile count fills the resident
// workgroup slots (zupdate_slots) evenly.
int zupdate_shape_count() { return 12; }
int zupdate_seg(int shape) { return shape % 3; }

int zupdate_slots(int prec, int shape, int cus) {
    const int v = shape / 3;
    if (prec == 0) {
        switch (v) {
            case 1: return zupdate_slots_t<double, 1, 1, 0>(cus);
            case 2: return zupdate_slots_t<double, 0, 1, 1>(cus);
            case 3: return zupdate_slots_t<double, 2, 1, 1>(cus);
            default: return zupdate_slots_t<double, 0, 1, 0>(cus);
        }
    }
    switch (v) {
        case 1: return zupdate_slots_t<float, 2, 1, 1>(cus);
        case 2: return zupdate_slots_t<float, 0, 3, 1>(cus);
        case 3: return zupdate_slots_t<float, 2, 3, 1>(cus);
        default: return zupdate_slots_t<float, 0, 1, 0>(cus);
    }
}

template <typename T>
static hipError_t launch_zupdate_p(const LaunchCtx& L, int shape, const GridDev* grids,
                                   const Tile* tiles, int ntiles, int stage, int do_cfl) {
    constexpr bool F = sizeof(T) == 4;
    switch (shape / 3) {
        case 1: return F ? launch_zupdate_t<T, 2, 1, 1>(L, grids, tiles, ntiles, stage, do_cfl)
                         : launch_zupdate_t<T, 1, 1, 0>(L, grids, tiles, ntiles, stage, do_cfl);
        case 2: return launch_zupdate_t<T, 0, F ? 3 : 1, 1>(L, grids, tiles, ntiles, stage, do_cfl);
        case 3: return launch_zupdate_t<T, 2, F ? 3 : 1, 1>(L, grids, tiles, ntiles, stage, do_cfl);
        default: return launch_zupdate_t<T, 0, 1, 0>(L, grids, tiles, ntiles, stage, do_cfl);
    }
}

hipError_t launch_zupdate(const LaunchCtx& L, int prec, int shape, const GridDev* grids,
                          const Tile* tiles, int ntiles, int stage, int do_cfl) {
    if (ntiles <= 0) return hipSuccess;
    if (prec == 0) return launch_zupdate_p<double>(L, shape, grids, tiles, ntiles, stage, do_cfl);
    return launch_zupdate_p<float>(L, shape, grids, tiles, ntiles, stage, do_cfl);
}

hipError_t launch_cfl(const LaunchCtx& L, const GridDev* grids, const Tile* tiles, int ntiles) {
    if (ntiles <= 0) return hipSuccess;
    const void* fn = (const void*)&k_cfl<UPDATE_KZ>;
    const int g = occupancy_grid(fn, 256, 0, ntiles, L.cus);
    hipLaunchKernelGGL((k_cfl<UPDATE_KZ>), dim3(g), dim3(256), 0, L.stream, grids, tiles, ntiles);
    return hipGetLastError();
}

static int flat_blocks(int64_t n, int cus) {
    int64_t b = (n + 255) / 256;
    const int64_t cap = (int64_t)cus * 16;
    if (b > cap) b = cap;
    return (int)(b < 1 ? 1 : b);
}

hipError_t launch_ghost_fill(const LaunchCtx& L, const GridDev* grids, const int64_t* off,
                             int ngrids, int64_t total, int copy_un) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ghost_fill, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, grids,
                       off, ngrids, copy_un);
    return hipGetLastError();
}

hipError_t launch_ghost_combine(const LaunchCtx& L, const GridDev* grids, const int64_t* off,
                                int ngrids, int64_t total, int stage) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ghost_combine, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream,
                       grids, off, ngrids, stage);
    return hipGetLastError();
}

hipError_t launch_floors(const LaunchCtx& L, const GridDev* grids, const int64_t* off, int ngrids,
                         int64_t total, int copy_un) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_floors, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, grids,
                       off, ngrids, copy_un);
    return hipGetLastError();
}

hipError_t launch_dt(const LaunchCtx& L, const GridDev* grids, int ngrids) {
    hipLaunchKernelGGL(k_dt, dim3((ngrids + 63) / 64), dim3(64), 0, L.stream, grids, ngrids);
    return hipGetLastError();
}

hipError_t launch_advance_t(const LaunchCtx& L, const GridDev* grids, int ngrids) {
    hipLaunchKernelGGL(k_advance_t, dim3((ngrids + 63) / 64), dim3(64), 0, L.stream, grids, ngrids);
    return hipGetLastError();
}

hipError_t launch_set_scal(const LaunchCtx& L, double* scal, int idx, double v) {
    hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(1), 0, L.stream, scal, idx, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64(const LaunchCtx& L, uint64_t* p, int64_t n, uint64_t v) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const LaunchCtx& L, const GridDev& G, const double* aos, double* dst) {
    const int64_t n = (int64_t)(G.nx + 2 * NG) * G.Ny * G.Nz;
    hipLaunchKernelGGL(k_aos_to_soa, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, aos, dst);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const LaunchCtx& L, const GridDev& G, const double* src, double* aos) {
    const int64_t n = (int64_t)(G.nx + 2 * NG) * G.Ny * G.Nz;
    hipLaunchKernelGGL(k_soa_to_aos, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, src, aos);
    return hipGetLastError();
}

hipError_t launch_flux_export(const LaunchCtx& L, const GridDev& G, int dir, double* out) {
    const int64_t n = (int64_t)(G.nx + 2 * NG + (dir == 0)) * (G.Ny + (dir == 1)) * (G.Nz + (dir == 2));
    if (G.prec == 0)
        hipLaunchKernelGGL(k_flux_export<double>, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, dir, out);
    else
        hipLaunchKernelGGL(k_flux_export<float>, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, dir, out);
    return hipGetLastError();
}

hipError_t launch_sg_force(const LaunchCtx& L, const GridDev& G, double* dU) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    hipLaunchKernelGGL(k_sg_force, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, dU);
    return hipGetLastError();
}

hipError_t launch_src(const LaunchCtx& L, const GridDev& G, double* dU, int which) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    hipLaunchKernelGGL(k_src, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, dU, which);
    return hipGetLastError();
}

hipError_t launch_combine_all(const LaunchCtx& L, const GridDev& G, int stage, const double* dU, double dt) {
    const int64_t n = (int64_t)(G.nx + 2 * NG) * G.Ny * G.Nz;
    hipLaunchKernelGGL(k_combine_all, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, G, stage, dU, dt);
    return hipGetLastError();
}

hipError_t launch_copy_state(const LaunchCtx& L, const GridDev& G, const double* src, double* dst) {
    hipLaunchKernelGGL(k_copy_state, dim3(flat_blocks(5 * G.fs, L.cus)), dim3(256), 0, L.stream, G, src, dst);
    return hipGetLastError();
}

// Fixed scratch of bsn_conserved_totals_noalloc: the block partial sums of
// one totals launch (TOTALS_BLOCKS x 5) and its five sums, in a file-scope
// device array.  Each device holds its own copy until the process ends; host
// code reaches it only through the address totals_scratch_setup returns
// (the array's name is a host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[TOTALS_SCRATCH_DOUBLES];

static std::mutex s_scratch_mu;
static double* s_scratch_ptr[BSN_MAX_DEVICES] = {};

// Once per device, with that device current: the device address of scratch.
hipError_t totals_scratch_setup(int device, double** ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= BSN_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(s_scratch_mu);
    if (!s_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        s_scratch_ptr[device] = (double*)p;
    }
    *ptr = s_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_totals(const LaunchCtx& L, const GridDev& G, double* part, int nblocks, double dV,
                         double* out) {
    hipLaunchKernelGGL(k_totals_partial, dim3(nblocks), dim3(256), 0, L.stream, G, part);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(64), 0, L.stream, part, nblocks, dV, out);
    return hipGetLastError();
}

hipError_t launch_prolong(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr, int64_t total,
                          const double* Uca, const double* Ucat, int all_cells, ProlongDst dst) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_prolong, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, grids, fmr,
                       Uca, Ucat, all_cells, dst);
    return hipGetLastError();
}

hipError_t launch_time_interp(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr,
                              int64_t total, double alpha) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_time_interp, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, grids,
                       fmr, alpha);
    return hipGetLastError();
}

hipError_t launch_restrict(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr, int64_t total,
                           int nfields) {
    hipLaunchKernelGGL(k_restrict, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, grids, fmr,
                       nfields);
    return hipGetLastError();
}

hipError_t launch_prolong_scalar(const LaunchCtx& L, const double* phc, const int nc[3], double* phf,
                                 const int nf[3], const int lo[3], int ratio) {
    const int64_t total = (int64_t)nf[0] * nf[1] * nf[2];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_prolong_scalar, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, phc,
                       nc[0], nc[1], nc[2], phf, nf[0], nf[1], nf[2], lo[0], lo[1], lo[2], ratio);
    return hipGetLastError();
}

hipError_t launch_scale(const LaunchCtx& L, double* a, int64_t n, double s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale, dim3(flat_blocks(n, L.cus)), dim3(256), 0, L.stream, a, n, s);
    return hipGetLastError();
}

hipError_t launch_fr_zero(const LaunchCtx& L, const FmrDev* fmr, int64_t total) {
    hipLaunchKernelGGL(k_fr_zero, dim3(flat_blocks(total, L.cus)), dim3(256), 0, L.stream, fmr);
    return hipGetLastError();
}

hipError_t launch_fr_coarse(const LaunchCtx& L, int prec, const GridDev* grids, const FmrDev* fmr,
                            int64_t nelem, double weight) {
    const dim3 grid(flat_blocks(nelem, L.cus), 5);     // one grid row per field
    if (prec == 0)
        hipLaunchKernelGGL(k_fr_coarse<double>, grid, dim3(256), 0, L.stream, grids, fmr, weight);
    else
        hipLaunchKernelGGL(k_fr_coarse<float>, grid, dim3(256), 0, L.stream, grids, fmr, weight);
    return hipGetLastError();
}

hipError_t launch_fr_fine(const LaunchCtx& L, int prec, const GridDev* grids, const FmrDev* fmr,
                          int64_t nelem, double weight) {
    const dim3 grid(flat_blocks(nelem, L.cus), 5);     // one grid row per field
    if (prec == 0)
        hipLaunchKernelGGL(k_fr_fine<double>, grid, dim3(256), 0, L.stream, grids, fmr, weight);
    else
        hipLaunchKernelGGL(k_fr_fine<float>, grid, dim3(256), 0, L.stream, grids, fmr, weight);
    return hipGetLastError();
}

hipError_t launch_bc_correct(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr, int64_t nelem) {
    hipLaunchKernelGGL(k_bc_correct, dim3(flat_blocks(nelem, L.cus)), dim3(256), 0, L.stream, grids, fmr);
    return hipGetLastError();
}

hipError_t launch_child_dt(const LaunchCtx& L, const GridDev* grids, int parent, int child, int ratio) {
    hipLaunchKernelGGL(k_child_dt, dim3(1), dim3(1), 0, L.stream, grids, parent, child, ratio);
    return hipGetLastError();
}

hipError_t launch_fold_dt(const LaunchCtx& L, const GridDev* grids, int root, int child, double factor) {
    hipLaunchKernelGGL(k_fold_dt, dim3(1), dim3(1), 0, L.stream, grids, root, child, factor);
    return hipGetLastError();
}

}  // namespace bsn
