// bsn_api.cpp -- host orchestration and the C ABI of libbsn_hip (include/bsn.h).
//
// A "plan" is a set of grids of one precision with device-resident tile
// work-lists for every kernel; a single grid, a batch of independent problems
// and each FMR level all run through plans.  One SSP-RK3 step of a plan is
//   stage 1: ghost fill (saving the raw state to Un) -> floors -> x/y/z fluxes
//            -> fused divergence + phi source + combine + floors
//   stage 2, 3: ghost fill -> fluxes -> fused update
//   then the stage-3 combination of ghost cells,
// which is src/euler3d.jl:220-260 with the RHS of src/euler3d.jl:182-205.  The
// start-of-stage floors of stages 2 and 3 are omitted: they follow the floors
// applied after the previous combination on the same active cells, and the
// floor is idempotent (a second application rewrites E to the same value).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <string>
#include <vector>
#include "bsn_internal.hpp"
#include "bsn_launch.hpp"
#include "bsn_poisson.hpp"
#include "bsn_bodies.hpp"
#include "bsn_slabs.hpp"
#include "../../include/bsn.h"

using namespace bsn;

static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return set_err(BSN_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                 \
                           hipGetErrorString(_e), __FILE__, __LINE__);                  \
    } while (0)

#define CHK(expr)                       \
    do {                                \
        int _r = (expr);                \
        if (_r != BSN_OK) return _r;    \
    } while (0)

static inline int64_t round_up(int64_t a, int64_t m) { return (a + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// Objects

struct bsn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 1;
    double* totals_scratch = nullptr;   // device address of the fixed totals scratch
    char name[256] = {0};
    LaunchCtx L() const { return LaunchCtx{stream, cus}; }
};

struct Plan;

struct bsn_grid {
    bsn_ctx* ctx = nullptr;
    GridDev h{};               // host copy of the descriptor
    bool params_set = false;
    double dx = 1, dy = 1, dz = 1;
    double* d_phi = nullptr;
    double* d_scratch = nullptr;   // dU-sized scratch (U layout)
    double* d_stage = nullptr;     // reference-layout staging
    size_t stage_bytes = 0;
    double* d_part = nullptr;      // totals partial sums
    double* d_work = nullptr;      // third state buffer of the out-of-place step
    double* d_work2 = nullptr;     // fourth state buffer (the fused FMR level step)
    SrcDev hs{};                   // point masses and sinks (host copy)
    SrcDev* d_src = nullptr;
    PoissonCtx* pois = nullptr;    // FFT Poisson solver (created on first use)
    double* d_rho_in = nullptr;    // density staging for bsn_solve_poisson_isolated
    bool self_grav = false;        // phi solved from each stage's density
    double* d_body = nullptr;      // body/diagnostic reductions: partial sums, then the sums
    double* h_pin = nullptr;       // two pinned slab buffers for file streaming
    size_t pin_bytes = 0;          // bytes of each
    Plan* plan = nullptr;          // single-grid plan
};

struct Plan {
    bsn_ctx* ctx = nullptr;
    std::vector<bsn_grid*> grids;
    int prec = 0;
    GridDev* d_grids = nullptr;
    int shape[3] = {0, 0, 0};
    Tile* d_ft[3] = {nullptr, nullptr, nullptr};
    int nft[3] = {0, 0, 0};
    Tile* d_ut = nullptr;
    int nut = 0;
    // out-of-place step: descriptor tables of stages 1..3 and z-update tiles
    GridDev* d_st[3] = {nullptr, nullptr, nullptr};
    Tile* d_zt = nullptr;
    int nzt = 0;
    int zshape = 0;
    int inplace = 0;     // 1: use the in-place step with materialised z fluxes
    int64_t* d_goff = nullptr;
    int64_t gtotal = 0;
    int64_t* d_aoff = nullptr;
    int64_t atotal = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    // profiling events (optional)
    bool prof = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_kind;   // 0..2 flux x/y/z, 3 update
};

struct bsn_batch {
    Plan* plan = nullptr;
};

struct bsn_fmr {
    bsn_ctx* ctx = nullptr;
    int nlev = 0;
    std::vector<bsn_grid*> lev;
    std::vector<Plan*> plans;
    GridDev* d_all = nullptr;
    std::vector<FmrDev> hf;
    FmrDev* d_f = nullptr;
    std::vector<double> alpha;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int inplace = 0;                 // 1: levels step in place with all fluxes materialised
    std::vector<GridDev*> d_fst;     // per level, the three stage tables of the fused step
    // self-gravity (fmr3d_step! with self_gravity = true, src/fmr3d.jl:451-473):
    // the frozen per-level phi of the root step, in buffers of its own
    int sg_on = 0;                   // bsn_fmr_run solves it every root step
    double sg_G = 1.0;
    std::vector<double*> d_sgphi;
};

// ---------------------------------------------------------------------------
// Tile lists

// Column tiles of a grid's (x, cross) plane for the y/z walks and the fused
// kernel: workgroups of 64 x 4 (x, rows) and, where a grid's rows end 32
// cells or fewer into their last x-tile, half-width workgroups of 32 x 8 for
// that tile (COL_HALF), so that the lanes past the row end do not idle.
static int col_tile_count(int nx, int ncross) {
    const int rem = nx % 64, nfull = nx / 64 + (rem > 32 ? 1 : 0);
    return nfull * ((ncross + 3) / 4) + (rem > 0 && rem <= 32 ? (ncross + 7) / 8 : 0);
}

template <typename F>
static void col_tiles(int nx, int ncross, F&& emit) {
    const int rem = nx % 64, nfull = nx / 64 + (rem > 32 ? 1 : 0);
    for (int b = 0; b < ncross; b += 4)
        for (int t = 0; t < nfull; ++t) emit(64 * t, b);
    if (rem > 0 && rem <= 32)
        for (int b = 0; b < ncross; b += 8) emit((64 * (nx / 64)) | COL_HALF, b);
}

// slots: workgroups the sweep kernel keeps resident (the persistent grid);
// the y and z walks split each column into runs of one balanced length so
// that the tile count is a whole multiple of the slots.
// Reorders a tile list of uneven tiles so that the persistent launch's
// uniform chunks (workgroup b of G takes tiles [b n / G, (b + 1) n / G), see
// chunk()) carry about equal work.  A batch of uneven grids gives runs of
// different lengths, and each grid's tiles sit together, so equal tile
// counts left some workgroups with several times the work of others.  Tiles
// go, longest first, to the least loaded chunk that still has room (chunk
// sizes fixed by chunk()); each chunk keeps its tiles in list order.  G is
// the launch's grid, min(slots, n).  Lists whose tiles weigh within 10 % of
// each other keep their spatial order.
template <typename Weight>
static void balance_chunks(std::vector<Tile>& t, int slots, Weight weight) {
    const int n = (int)t.size();
    const int G = std::min(slots, n);
    if (G <= 1) return;
    std::vector<double> w(n);
    double wmin = 1e300, wmax = 0.0;
    for (int i = 0; i < n; ++i) {
        w[i] = weight(t[i]);
        wmin = std::min(wmin, w[i]);
        wmax = std::max(wmax, w[i]);
    }
    if (wmax <= 1.1 * wmin) return;
    std::vector<int> cap(G);
    for (int b = 0; b < G; ++b)
        cap[b] = (int)(((int64_t)(b + 1) * n) / G) - (int)(((int64_t)b * n) / G);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
    using Load = std::pair<double, int>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int b = 0; b < G; ++b)
        if (cap[b] > 0) heap.push({0.0, b});
    std::vector<std::vector<int>> lists(G);
    for (int i : order) {
        const Load top = heap.top();
        heap.pop();
        const int b = top.second;
        lists[b].push_back(i);
        if ((int)lists[b].size() < cap[b]) heap.push({top.first + w[i], b});
    }
    std::vector<Tile> out;
    out.reserve(n);
    for (auto& l : lists) {
        std::sort(l.begin(), l.end());
        for (int i : l) out.push_back(t[i]);
    }
    t.swap(out);
}

static std::vector<Tile> flux_tiles(const std::vector<GridDev>& G, int dir, int shape, int slots) {
    int W, H;
    flux_shape_dims(dir, shape, W, H);
    std::vector<Tile> t;
    if (dir == 0 && flux_walk_len(0, shape) > 0) {
        // x walks: 4 rows per tile, interfaces 0 .. nx-1 of a row in full
        // tiles of 64 K (K per lane) and one for the rest; edge tiles take
        // the last interface of 256 rows each
        const int per = 64 * flux_walk_len(0, shape);
        for (int g = 0; g < (int)G.size(); ++g) {
            const GridDev& d = G[g];
            for (int k = 0; k < d.nz; ++k)
                for (int j0 = 0; j0 < d.ny; j0 += 4)
                    for (int a = 0; a < d.nx; a += per) {
                        const int n = d.nx - a < per ? d.nx - a : per;
                        t.push_back({g, a, j0, k | (n << ZT_LEN_SHIFT)});
                    }
            for (int r0 = 0; r0 < d.ny * d.nz; r0 += 256) t.push_back({g, -1, r0, 0});
        }
        return t;
    }
    if (dir != 0 && flux_walk_len(dir, shape) > 0) {
        // register walks: column tiles (col_tiles), each column split into
        // runs along dir.
        // The persistent launch hands every workgroup the same number of
        // tiles, so a tile count just above a multiple of the slots leaves
        // most workgroups idle for a final pass (2304 runs of 32 over 1024
        // slots take three passes).  The runs of all grids therefore share
        // one length Lt <= 1.1 * Lmax, chosen so that the column-interfaces
        // divide evenly over m * slots runs.
        const int Lmax = flux_walk_len(dir, shape);
        double work = 0;    // column-interfaces
        for (const GridDev& d : G)
            work += (double)col_tile_count(d.nx, dir == 1 ? d.nz : d.ny) *
                    ((dir == 1 ? d.ny : d.nz) + 1);
        const double sl = slots > 0 ? slots : 1;
        int m = (int)std::ceil(work / (sl * Lmax * 1.1));
        if (m < 1) m = 1;
        // a batch too small to fill the slots keeps runs of at least 8:
        // shorter runs would only add window starts
        const double Lt = std::max(work / (m * sl), 8.0);
        for (int g = 0; g < (int)G.size(); ++g) {
            const GridDev& d = G[g];
            const int nf = (dir == 1 ? d.ny : d.nz) + 1;      // interfaces along dir
            const int ncross = dir == 1 ? d.nz : d.ny;
            int nr = (int)(nf / Lt + 0.5);
            if (nr < 1) nr = 1;
            while ((nf + nr - 1) / nr > ZT_MAX_LEN) ++nr;
            for (int r = 0; r < nr; ++r) {
                const int f0 = (int)((int64_t)nf * r / nr), f1 = (int)((int64_t)nf * (r + 1) / nr);
                if (f1 <= f0) continue;
                col_tiles(d.nx, ncross, [&](int a, int b) {
                    t.push_back({g, a, b, f0 | ((f1 - f0) << ZT_LEN_SHIFT)});
                });
            }
        }
        // a run of L interfaces reads L + 5 cells
        balance_chunks(t, slots, [](const Tile& x) { return (double)((x.c >> ZT_LEN_SHIFT) + 5); });
        return t;
    }
    for (int g = 0; g < (int)G.size(); ++g) {
        const GridDev& d = G[g];
        if (dir == 0) {
            for (int k = 0; k < d.nz; ++k) {
                for (int j0 = 0; j0 < d.ny; j0 += H)
                    for (int f0 = 0; f0 < d.nx; f0 += W) t.push_back({g, f0, j0, k});
                for (int j0 = 0; j0 < d.ny; j0 += W * H) t.push_back({g, -1, j0, k});
            }
        } else if (dir == 1) {
            for (int k = 0; k < d.nz; ++k)
                for (int i0 = 0; i0 < d.nx; i0 += W)
                    for (int f0 = 0; f0 <= d.ny; f0 += H) t.push_back({g, f0, i0, k});
        } else {
            for (int j = 0; j < d.ny; ++j)
                for (int i0 = 0; i0 < d.nx; i0 += W)
                    for (int f0 = 0; f0 <= d.nz; f0 += H) t.push_back({g, f0, i0, j});
        }
    }
    return t;
}

static std::vector<Tile> update_tiles(const std::vector<GridDev>& G) {
    std::vector<Tile> t;
    for (int g = 0; g < (int)G.size(); ++g) {
        const GridDev& d = G[g];
        for (int k0 = 0; k0 < d.nz; k0 += UPDATE_KZ)
            for (int j0 = 0; j0 < d.ny; j0 += 4)
                for (int i0 = 0; i0 < d.nx; i0 += 64) t.push_back({g, i0, j0, k0});
    }
    return t;
}

// Column tiles of the fused z-sweep + update kernel: 64 x 4 columns of one
// grid and a run of cells along z.  seg 0/1: runs of 16/32 cells.  seg 2: the
// runs of each grid have one length (split as evenly as the grid's nz allows),
// chosen so that the tile count is close to a whole multiple of the slots
// (resident workgroups) and no run exceeds 96 cells: the persistent launch
// then gives every workgroup the same work, and long runs amortise the
// five-cell window start-up.
static std::vector<Tile> zupdate_tiles(const std::vector<GridDev>& G, int seg, int slots) {
    std::vector<Tile> t;
    std::vector<int> nseg(G.size(), 1);
    if (seg == 2) {
        double work = 0;    // column-cells
        for (const GridDev& d : G) work += (double)col_tile_count(d.nx, d.ny) * d.nz;
        const double s = slots > 0 ? slots : 1;
        int m = 1;
        while (work / (m * s) > 96.0) ++m;
        const double Lt = work / (m * s);
        for (size_t g = 0; g < G.size(); ++g) {
            int n = (int)(G[g].nz / Lt + 0.5);
            nseg[g] = n < 1 ? 1 : n;
        }
    } else {
        const int kz = seg == 1 ? 32 : 16;
        for (size_t g = 0; g < G.size(); ++g) nseg[g] = (G[g].nz + kz - 1) / kz;
    }
    for (int g = 0; g < (int)G.size(); ++g) {
        const GridDev& d = G[g];
        int n = nseg[g];
        while ((d.nz + n - 1) / n > ZT_MAX_LEN) ++n;
        for (int s = 0; s < n; ++s) {
            const int k0 = (int)((int64_t)d.nz * s / n), k1 = (int)((int64_t)d.nz * (s + 1) / n);
            if (k1 <= k0) continue;
            const int c = k0 | ((k1 - k0) << ZT_LEN_SHIFT);
            col_tiles(d.nx, d.ny, [&](int a, int b) { t.push_back({g, a, b, c}); });
        }
    }
    // a run of L cells reads a window of L + 5
    balance_chunks(t, slots, [](const Tile& x) { return (double)((x.c >> ZT_LEN_SHIFT) + 5); });
    return t;
}

template <typename T>
static int upload_vec(T** dptr, const std::vector<T>& v) {
    if (*dptr) { HIPCHK(hipFree(*dptr)); *dptr = nullptr; }
    if (v.empty()) return BSN_OK;
    HIPCHK(hipMalloc((void**)dptr, v.size() * sizeof(T)));
    HIPCHK(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return BSN_OK;
}

static std::vector<GridDev> host_descs(const Plan& P) {
    std::vector<GridDev> v;
    for (auto* g : P.grids) v.push_back(g->h);
    return v;
}

static void graph_drop(hipGraphExec_t& e, hipGraph_t& g) {
    if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    if (g) { (void)hipGraphDestroy(g); g = nullptr; }
}

static int plan_set_shape(Plan& P, int dir, int shape) {
    auto G = host_descs(P);
    auto t = flux_tiles(G, dir, shape, flux_slots(dir, P.prec, shape, P.ctx->cus));
    CHK(upload_vec(&P.d_ft[dir], t));
    P.nft[dir] = (int)t.size();
    P.shape[dir] = shape;
    graph_drop(P.gexec, P.graph);
    return BSN_OK;
}

static int ensure_work(bsn_grid* g) {
    if (g->d_work) return BSN_OK;
    HIPCHK(hipMalloc((void**)&g->d_work, 5 * g->h.fs * 8));
    HIPCHK(hipMemset(g->d_work, 0, 5 * g->h.fs * 8));
    return BSN_OK;
}

static int ensure_work2(bsn_grid* g) {
    if (g->d_work2) return BSN_OK;
    HIPCHK(hipMalloc((void**)&g->d_work2, 5 * g->h.fs * 8));
    HIPCHK(hipMemset(g->d_work2, 0, 5 * g->h.fs * 8));
    return BSN_OK;
}

// Descriptor tables.  In-place table: U = Uout = the state, Un = base copy.
// Out-of-place step with the state A, B = the Un buffer and C = the work
// buffer:  stage 1 reads A, writes B;  stage 2 reads B, writes C;  stage 3
// reads C, writes A.  A stays the raw step-start state until stage 3
// overwrites it cell by cell, so it is also the base state Un.  Every stage
// reads its input with ghost cells mapped to active cells on load, and stage
// 1 floors active cells on load, so no ghost fill or floor pass runs and A is
// never modified before stage 3.
static int plan_refresh(Plan& P) {
    HIPCHK(hipSetDevice(P.ctx->device));
    auto G = host_descs(P);
    const size_t bytes = G.size() * sizeof(GridDev);
    if (!P.d_grids) HIPCHK(hipMalloc((void**)&P.d_grids, bytes));
    HIPCHK(hipMemcpy(P.d_grids, G.data(), bytes, hipMemcpyHostToDevice));
    std::vector<GridDev> st[3];
    for (size_t gi = 0; gi < G.size(); ++gi) {
        bsn_grid* g = P.grids[gi];
        CHK(ensure_work(g));
        double* A = g->h.U;
        double* B = g->h.Un;
        double* C = g->d_work;
        GridDev d = G[gi];
        d.Un = A;
        d.lmode = 3; d.U = A; d.Uout = B; st[0].push_back(d);
        d.lmode = 1; d.U = B; d.Uout = C; st[1].push_back(d);
        d.lmode = 1; d.U = C; d.Uout = A; st[2].push_back(d);
    }
    for (int s = 0; s < 3; ++s) {
        if (!P.d_st[s]) HIPCHK(hipMalloc((void**)&P.d_st[s], bytes));
        HIPCHK(hipMemcpy(P.d_st[s], st[s].data(), bytes, hipMemcpyHostToDevice));
    }
    return BSN_OK;
}

static int plan_set_zshape(Plan& P, int shape) {
    auto G = host_descs(P);
    auto t = zupdate_tiles(G, zupdate_seg(shape), zupdate_slots(P.prec, shape, P.ctx->cus));
    CHK(upload_vec(&P.d_zt, t));
    P.nzt = (int)t.size();
    P.zshape = shape;
    graph_drop(P.gexec, P.graph);
    return BSN_OK;
}

static int plan_build(Plan& P) {
    HIPCHK(hipSetDevice(P.ctx->device));
    CHK(plan_refresh(P));
    auto G = host_descs(P);
    for (int d = 0; d < 3; ++d) CHK(plan_set_shape(P, d, P.shape[d]));
    CHK(plan_set_zshape(P, P.zshape));
    auto ut = update_tiles(G);
    CHK(upload_vec(&P.d_ut, ut));
    P.nut = (int)ut.size();
    std::vector<int64_t> go(G.size() + 1, 0), ao(G.size() + 1, 0);
    for (size_t g = 0; g < G.size(); ++g) {
        const GridDev& d = G[g];
        go[g + 1] = go[g] + 2 * BSN_NG * ((int64_t)d.Ny * d.Nz + (int64_t)d.nx * d.Nz + (int64_t)d.nx * d.ny);
        ao[g + 1] = ao[g] + (int64_t)d.nx * d.ny * d.nz;
    }
    CHK(upload_vec(&P.d_goff, go));
    CHK(upload_vec(&P.d_aoff, ao));
    P.gtotal = go.back();
    P.atotal = ao.back();
    return BSN_OK;
}

static void plan_free(Plan* P) {
    if (!P) return;
    (void)hipSetDevice(P->ctx->device);
    graph_drop(P->gexec, P->graph);
    for (auto e : P->ev) (void)hipEventDestroy(e);
    (void)hipFree(P->d_grids);
    for (int d = 0; d < 3; ++d) (void)hipFree(P->d_ft[d]);
    (void)hipFree(P->d_ut);
    for (int s = 0; s < 3; ++s) (void)hipFree(P->d_st[s]);
    (void)hipFree(P->d_zt);
    (void)hipFree(P->d_goff);
    (void)hipFree(P->d_aoff);
    delete P;
}

static int plan_new(bsn_ctx* ctx, const std::vector<bsn_grid*>& grids, Plan** out) {
    if (grids.empty()) return set_err(BSN_ERR_ARG, "plan with no grids");
    Plan* P = new Plan();
    P->ctx = ctx;
    P->grids = grids;
    P->prec = grids[0]->h.prec;
    for (auto* g : grids) {
        if (g->h.prec != P->prec) { delete P; return set_err(BSN_ERR_ARG, "grids of mixed precision"); }
        if (g->ctx != ctx) { delete P; return set_err(BSN_ERR_ARG, "grids of different contexts"); }
    }
    // starting shapes (kept until bsn_batch_autotune or set_*shape*): y and z
    // register walks in balanced runs (the autotuner's choice from 32^3
    // batches to the 256^3 box), the fused kernel with the LDS ring, the
    // early load and balanced runs; the x sweep tiled (64x4 split, 16x16
    // FP64), which unlike the x walks keeps its lanes busy on narrow grids.
    if (P->prec == 1) { P->shape[0] = 0; P->shape[1] = FLUX_WALK_SHAPE0 + 2; P->shape[2] = FLUX_WALK_SHAPE0 + 2; P->zshape = 11; }
    else              { P->shape[0] = 2; P->shape[1] = FLUX_WALK_SHAPE0 + 2; P->shape[2] = FLUX_WALK_SHAPE0 + 2; P->zshape = 11; }
    int r = plan_build(*P);
    if (r != BSN_OK) { plan_free(P); return r; }
    *out = P;
    return BSN_OK;
}

// ---------------------------------------------------------------------------
// Plan stages

static void prof_mark(Plan& P, int kind) {
    if (!P.prof) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, P.ctx->stream);
    P.ev.push_back(e);
    P.ev_kind.push_back(kind);
}

static int plan_fluxes(Plan& P, const GridDev* table = nullptr, int ndirs = 3) {
    const LaunchCtx L = P.ctx->L();
    if (!table) table = P.d_grids;
    for (int d = 0; d < ndirs; ++d) {
        prof_mark(P, d);
        HIPCHK(launch_flux(L, d, P.prec, P.shape[d], table, P.d_ft[d], P.nft[d], 0));
    }
    prof_mark(P, 9);
    return BSN_OK;
}

static int plan_update(Plan& P, int stage, int do_cfl, double* dU_out = nullptr) {
    prof_mark(P, 3);
    HIPCHK(launch_update(P.ctx->L(), P.prec, P.d_grids, P.d_ut, P.nut, stage, do_cfl, dU_out));
    prof_mark(P, 9);
    return BSN_OK;
}

static int plan_ghost_fill(Plan& P, int copy_un, const GridDev* table = nullptr) {
    HIPCHK(launch_ghost_fill(P.ctx->L(), table ? table : P.d_grids, P.d_goff, (int)P.grids.size(),
                             P.gtotal, copy_un));
    return BSN_OK;
}

static int plan_floors(Plan& P, int copy_un) {
    HIPCHK(launch_floors(P.ctx->L(), P.d_grids, P.d_aoff, (int)P.grids.size(), P.atotal, copy_un));
    return BSN_OK;
}

static int plan_ghost_combine(Plan& P, int stage, const GridDev* table = nullptr) {
    HIPCHK(launch_ghost_combine(P.ctx->L(), table ? table : P.d_grids, P.d_goff,
                                (int)P.grids.size(), P.gtotal, stage));
    return BSN_OK;
}

static int plan_step_inplace(Plan& P, int fused_cfl);

// One SSP-RK3 step, out of place (see plan_refresh): per stage the x and y
// sweeps write their fluxes and one fused kernel computes the z fluxes and
// finishes every cell.  Bit-identical to the in-place step below.
static int plan_step(Plan& P, int fused_cfl) {
    if (P.inplace) return plan_step_inplace(P, fused_cfl);
    const LaunchCtx L = P.ctx->L();
    for (int s = 1; s <= 3; ++s) {
        const GridDev* T = P.d_st[s - 1];
        CHK(plan_fluxes(P, T, 2));
        for (auto* g : P.grids) {
            if (!g->self_grav) continue;
            GridDev in = g->h;
            in.U = s == 1 ? g->h.U : s == 2 ? g->h.Un : g->d_work;
            HIPCHK(poisson_solve_state(*g->pois, in, s == 1, g->d_phi, L.stream));
        }
        prof_mark(P, 3);
        HIPCHK(launch_zupdate(L, P.prec, P.zshape, T, P.d_zt, P.nzt, s, fused_cfl && s == 3));
        prof_mark(P, 9);
    }
    CHK(plan_ghost_combine(P, 3, P.d_st[2]));
    return BSN_OK;
}

// The same step in place with all three flux arrays materialised (the form
// the FMR levels use, since their flux registers read the z faces).
static int plan_step_inplace(Plan& P, int fused_cfl) {
    for (int s = 1; s <= 3; ++s) {
        if (s == 1) {
            CHK(plan_ghost_fill(P, 1));
            CHK(plan_floors(P, 1));
        } else {
            CHK(plan_ghost_fill(P, 0));
        }
        CHK(plan_fluxes(P));
        for (auto* g : P.grids)
            if (g->self_grav)
                HIPCHK(poisson_solve_state(*g->pois, g->h, 0, g->d_phi, P.ctx->stream));
        CHK(plan_update(P, s, fused_cfl && s == 3));
    }
    CHK(plan_ghost_combine(P, 3));
    return BSN_OK;
}

static int plan_cfl(Plan& P, double cfl, double t_end) {
    const LaunchCtx L = P.ctx->L();
    for (auto* g : P.grids) {
        HIPCHK(launch_fill_u64(L, (uint64_t*)&g->h.scal[SCAL_SMAX], 1, 0));
        HIPCHK(launch_set_scal(L, g->h.scal, SCAL_CFL, cfl));
        HIPCHK(launch_set_scal(L, g->h.scal, SCAL_TEND, t_end));
    }
    HIPCHK(launch_cfl(L, P.d_grids, P.d_ut, P.nut));
    return BSN_OK;
}

// ---------------------------------------------------------------------------
// Library / context

extern "C" int bsn_version(void) { return 10000; }
extern "C" const char* bsn_last_error(void) { return g_err.c_str(); }

extern "C" int bsn_ctx_create(int device, bsn_ctx** out) {
    if (!out) return set_err(BSN_ERR_ARG, "out is NULL");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(BSN_ERR_ARG, "device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    bsn_ctx* c = new bsn_ctx();
    c->device = device;
    c->cus = prop.multiProcessorCount;
    snprintf(c->name, sizeof c->name, "%s (%s), %d CUs, %zu MB", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, (size_t)(prop.totalGlobalMem >> 20));
    hipError_t e = totals_scratch_setup(device, &c->totals_scratch);
    if (e != hipSuccess) {
        delete c;
        return set_err(BSN_ERR_HIP, "address of the totals scratch: %s", hipGetErrorString(e));
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return set_err(BSN_ERR_HIP, "stream: %s", hipGetErrorString(e)); }
    *out = c;
    return BSN_OK;
}

extern "C" int bsn_ctx_destroy(bsn_ctx* c) {
    if (!c) return BSN_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return BSN_OK;
}

extern "C" int bsn_ctx_synchronize(bsn_ctx* c) {
    if (!c) return set_err(BSN_ERR_ARG, "ctx is NULL");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BSN_OK;
}

extern "C" int bsn_ctx_info(bsn_ctx* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return set_err(BSN_ERR_ARG, "bad arguments");
    snprintf(buf, (size_t)len, "%s", c->name);
    return BSN_OK;
}

// ---------------------------------------------------------------------------
// Grids

extern "C" int bsn_grid_create(bsn_ctx* ctx, int64_t nx, int64_t ny, int64_t nz, int precision,
                               bsn_grid** out) {
    if (!ctx || !out) return set_err(BSN_ERR_ARG, "NULL argument");
    if (nx < 1 || ny < 1 || nz < 1 || nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20))
        return set_err(BSN_ERR_ARG, "bad grid dims %lld x %lld x %lld", (long long)nx,
                       (long long)ny, (long long)nz);
    if (precision != BSN_PREC_FP64 && precision != BSN_PREC_SPLIT)
        return set_err(BSN_ERR_ARG, "precision %d", precision);
    // the kernels address one z plane of a field (state or x/y faces)
    // through a 32-bit byte offset from the plane's base, below 2 GiB
    {
        const int64_t plane = (int64_t)(ny + 2 * BSN_NG + 1) * (BSN_XOFF + nx + 2 * BSN_NG + 16) * 8;
        if (plane >= ((int64_t)1 << 31))
            return set_err(BSN_ERR_ARG, "an x-y plane of %lld x %lld cells exceeds 2 GiB per field",
                           (long long)nx, (long long)ny);
    }
    HIPCHK(hipSetDevice(ctx->device));
    bsn_grid* g = new bsn_grid();
    g->ctx = ctx;
    GridDev& h = g->h;
    h.nx = (int)nx; h.ny = (int)ny; h.nz = (int)nz;
    h.Ny = h.ny + 2 * BSN_NG;
    h.Nz = h.nz + 2 * BSN_NG;
    const int Nx = h.nx + 2 * BSN_NG;
    h.pitch = (int)round_up(BSN_XOFF + Nx, 16);
    h.fs = round_up((int64_t)h.pitch * h.Ny * h.Nz, 64);
    h.prec = precision;
    const int64_t esz = precision == BSN_PREC_FP64 ? 8 : 4;
    h.fp[0] = (int)round_up(h.nx + 1, 32);
    h.fp[1] = (int)round_up(h.nx, 32);
    h.fp[2] = (int)round_up(h.nx, 32);
    h.ffs[0] = round_up((int64_t)h.fp[0] * h.ny * h.nz, 128);
    h.ffs[1] = round_up((int64_t)h.fp[1] * (h.ny + 1) * h.nz, 128);
    h.ffs[2] = round_up((int64_t)h.fp[2] * h.ny * (h.nz + 1), 128);
    h.gamma = 1.4; h.gfac = (h.gamma + 1.0) / (2.0 * h.gamma);
    h.gf32[0] = (float)h.gamma;
    h.gf32[1] = (float)h.gfac;
    h.gf32[2] = (float)(1.0 / (h.gamma - 1.0));
    h.idx = h.idy = h.idz = 1.0;
    h.bc = 0;
    auto fail = [&](hipError_t e) {
        (void)hipFree(h.U); (void)hipFree(h.Un); (void)hipFree(h.scal);
        for (int d = 0; d < 3; ++d) (void)hipFree(h.F[d]);
        delete g;
        return set_err(BSN_ERR_HIP, "allocation failed: %s", hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&h.U, 5 * h.fs * 8)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&h.Un, 5 * h.fs * 8)) != hipSuccess) return fail(e);
    for (int d = 0; d < 3; ++d)
        if ((e = hipMalloc(&h.F[d], 5 * h.ffs[d] * esz)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&h.scal, SCAL_COUNT * 8)) != hipSuccess) return fail(e);
    h.Uout = h.U;
    // padding and never-written flux entries stay defined
    (void)hipMemset(h.U, 0, 5 * h.fs * 8);
    (void)hipMemset(h.Un, 0, 5 * h.fs * 8);
    for (int d = 0; d < 3; ++d) (void)hipMemset(h.F[d], 0, 5 * h.ffs[d] * esz);
    (void)hipMemset(h.scal, 0, SCAL_COUNT * 8);
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e);
    *out = g;
    return BSN_OK;
}

extern "C" int bsn_grid_destroy(bsn_grid* g) {
    if (!g) return BSN_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    plan_free(g->plan);
    GridDev& h = g->h;
    (void)hipFree(h.U); (void)hipFree(h.Un); (void)hipFree(h.scal);
    for (int d = 0; d < 3; ++d) (void)hipFree(h.F[d]);
    (void)hipFree(g->d_phi);
    (void)hipFree(g->d_scratch);
    (void)hipFree(g->d_stage);
    (void)hipFree(g->d_part);
    (void)hipFree(g->d_work);
    (void)hipFree(g->d_work2);
    (void)hipFree(g->d_src);
    if (g->pois) { poisson_destroy(*g->pois); delete g->pois; }
    (void)hipFree(g->d_rho_in);
    (void)hipFree(g->d_body);
    if (g->h_pin) (void)hipHostFree(g->h_pin);
    delete g;
    return BSN_OK;
}

static int64_t ref_count(const GridDev& h) {
    return 5LL * (h.nx + 2 * BSN_NG) * h.Ny * h.Nz;
}

static int ensure_stage(bsn_grid* g, size_t bytes) {
    if (g->stage_bytes >= bytes) return BSN_OK;
    if (g->d_stage) HIPCHK(hipFree(g->d_stage));
    g->d_stage = nullptr;
    g->stage_bytes = 0;
    HIPCHK(hipMalloc((void**)&g->d_stage, bytes));
    g->stage_bytes = bytes;
    return BSN_OK;
}

static int ensure_scratch(bsn_grid* g) {
    if (g->d_scratch) return BSN_OK;
    HIPCHK(hipMalloc((void**)&g->d_scratch, 5 * g->h.fs * 8));
    HIPCHK(hipMemset(g->d_scratch, 0, 5 * g->h.fs * 8));
    return BSN_OK;
}

static int grid_plan(bsn_grid* g) {
    if (g->plan) return BSN_OK;
    return plan_new(g->ctx, {g}, &g->plan);
}

static int sync_ctx(bsn_ctx* c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    return BSN_OK;
}

// reference layout (host or device) -> device buffer with U's layout
static int import_ref(bsn_grid* g, const double* src, int where, double* dst) {
    const LaunchCtx L = g->ctx->L();
    const size_t bytes = (size_t)ref_count(g->h) * 8;
    const double* from = src;
    if (where == BSN_HOST) {
        CHK(ensure_stage(g, bytes));
        HIPCHK(hipMemcpyAsync(g->d_stage, src, bytes, hipMemcpyHostToDevice, g->ctx->stream));
        from = g->d_stage;
    }
    HIPCHK(launch_aos_to_soa(L, g->h, from, dst));
    return BSN_OK;
}

static int export_ref(bsn_grid* g, const double* src, double* dst, int where) {
    const LaunchCtx L = g->ctx->L();
    const size_t bytes = (size_t)ref_count(g->h) * 8;
    if (where == BSN_HOST) {
        CHK(ensure_stage(g, bytes));
        HIPCHK(launch_soa_to_aos(L, g->h, src, g->d_stage));
        HIPCHK(hipMemcpyAsync(dst, g->d_stage, bytes, hipMemcpyDeviceToHost, g->ctx->stream));
    } else {
        HIPCHK(launch_soa_to_aos(L, g->h, src, dst));
    }
    return BSN_OK;
}

extern "C" int bsn_grid_set_state(bsn_grid* g, const double* U, int where) {
    if (!g || !U) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(import_ref(g, U, where, g->h.U));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_grid_get_state(bsn_grid* g, double* U, int where) {
    if (!g || !U) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(export_ref(g, g->h.U, U, where));
    return sync_ctx(g->ctx);
}

static int grid_changed(bsn_grid* g) {
    // propagate descriptor changes to every plan that holds this grid
    if (g->plan) CHK(plan_refresh(*g->plan));
    return BSN_OK;
}

// Uploads the grid's point masses and sinks (with its current cell sizes)
// and points the descriptor at them, or clears the pointer when none are set.
static int src_upload(bsn_grid* g) {
    SrcDev& S = g->hs;
    S.dx = g->dx; S.dy = g->dy; S.dz = g->dz;
    const bool any = S.nbh > 0 || S.nsink > 0;
    if (any) {
        if (!g->d_src) HIPCHK(hipMalloc((void**)&g->d_src, sizeof(SrcDev)));
        HIPCHK(hipMemcpy(g->d_src, &S, sizeof(SrcDev), hipMemcpyHostToDevice));
    }
    const SrcDev* want = any ? g->d_src : nullptr;
    if (g->h.src != want) {
        g->h.src = want;
        CHK(grid_changed(g));
    }
    return BSN_OK;
}

static int fill_masses(SrcDev& S, int n, const double* pos, const double* mass, const double* eps,
                       double x0, double y0, double z0) {
    if (n < 0 || n > BSN_MAX_BH) return set_err(BSN_ERR_ARG, "point mass count %d (max %d)", n, BSN_MAX_BH);
    if (n > 0 && (!pos || !mass || !eps)) return set_err(BSN_ERR_ARG, "NULL point-mass array");
    S.nbh = n;
    S.gx0[0] = x0; S.gx0[1] = y0; S.gx0[2] = z0;
    for (int b = 0; b < n; ++b) {
        S.px[b] = pos[3 * b]; S.py[b] = pos[3 * b + 1]; S.pz[b] = pos[3 * b + 2];
        S.mass[b] = mass[b]; S.eps[b] = eps[b];
    }
    return BSN_OK;
}

static int fill_sinks(SrcDev& S, int n, const double* pos, const double* vel, const double* r_sink,
                      const double* t_sink, int torque_free, double rho_sink_min, double x0,
                      double y0, double z0) {
    if (n < 0 || n > BSN_MAX_BH) return set_err(BSN_ERR_ARG, "sink count %d (max %d)", n, BSN_MAX_BH);
    if (n > 0 && (!pos || !vel || !r_sink || !t_sink)) return set_err(BSN_ERR_ARG, "NULL sink array");
    S.nsink = n;
    S.torque_free = torque_free ? 1 : 0;
    S.rho_sink_min = rho_sink_min;
    S.sx0[0] = x0; S.sx0[1] = y0; S.sx0[2] = z0;
    for (int b = 0; b < n; ++b) {
        S.spx[b] = pos[3 * b]; S.spy[b] = pos[3 * b + 1]; S.spz[b] = pos[3 * b + 2];
        S.svx[b] = vel[3 * b]; S.svy[b] = vel[3 * b + 1]; S.svz[b] = vel[3 * b + 2];
        S.rs[b] = r_sink[b]; S.ts[b] = t_sink[b];
    }
    return BSN_OK;
}

extern "C" int bsn_grid_set_point_masses(bsn_grid* g, int n, const double* pos, const double* mass,
                                         const double* eps, double x0, double y0, double z0) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    SrcDev S = g->hs;
    CHK(fill_masses(S, n, pos, mass, eps, x0, y0, z0));
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(hipStreamSynchronize(g->ctx->stream));
    g->hs = S;
    return src_upload(g);
}

extern "C" int bsn_grid_set_sinks(bsn_grid* g, int n, const double* pos, const double* vel,
                                  const double* r_sink, const double* t_sink, int torque_free,
                                  double rho_sink_min, double x0, double y0, double z0) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    SrcDev S = g->hs;
    CHK(fill_sinks(S, n, pos, vel, r_sink, t_sink, torque_free, rho_sink_min, x0, y0, z0));
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(hipStreamSynchronize(g->ctx->stream));
    g->hs = S;
    return src_upload(g);
}

// One source on a caller's dU from the grid's state, with the bodies of the
// call (nothing is kept on the grid).
static int add_src(bsn_grid* g, double* dU, int where, double dx, double dy, double dz,
                   const SrcDev& S0, int which) {
    if (!g || !dU) return set_err(BSN_ERR_ARG, "NULL argument");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return set_err(BSN_ERR_ARG, "cell sizes must be > 0");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    SrcDev S = S0;
    S.dx = dx; S.dy = dy; S.dz = dz;
    CHK(ensure_scratch(g));
    SrcDev* d = nullptr;
    HIPCHK(hipMalloc((void**)&d, sizeof(SrcDev)));
    hipError_t e = hipMemcpy(d, &S, sizeof(SrcDev), hipMemcpyHostToDevice);
    int rc = BSN_OK;
    if (e != hipSuccess) rc = set_err(BSN_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
    GridDev h = g->h;
    h.src = d;
    if (rc == BSN_OK) rc = import_ref(g, dU, where, g->d_scratch);
    if (rc == BSN_OK) {
        e = launch_src(g->ctx->L(), h, g->d_scratch, which);
        if (e != hipSuccess) rc = set_err(BSN_ERR_HIP, "launch: %s", hipGetErrorString(e));
    }
    if (rc == BSN_OK) rc = export_ref(g, g->d_scratch, dU, where);
    if (rc == BSN_OK) rc = sync_ctx(g->ctx);
    (void)hipStreamSynchronize(g->ctx->stream);
    (void)hipFree(d);
    return rc;
}

extern "C" int bsn_add_bh_gravity_source(bsn_grid* g, double* dU, int where, double dx, double dy,
                                         double dz, int n, const double* pos, const double* mass,
                                         const double* eps, double x0, double y0, double z0) {
    SrcDev S{};
    CHK(fill_masses(S, n, pos, mass, eps, x0, y0, z0));
    return add_src(g, dU, where, dx, dy, dz, S, 0);
}

extern "C" int bsn_add_sink_sources(bsn_grid* g, double* dU, int where, double dx, double dy,
                                    double dz, int n, const double* pos, const double* vel,
                                    const double* r_sink, const double* t_sink, int torque_free,
                                    double rho_sink_min, double x0, double y0, double z0) {
    SrcDev S{};
    CHK(fill_sinks(S, n, pos, vel, r_sink, t_sink, torque_free, rho_sink_min, x0, y0, z0));
    return add_src(g, dU, where, dx, dy, dz, S, 1);
}

static int ensure_poisson(bsn_grid* g, double dx, double dy, double dz) {
    PoissonCtx* P = g->pois;
    if (P && P->dx == dx && P->dy == dy && P->dz == dz) return BSN_OK;
    HIPCHK(hipStreamSynchronize(g->ctx->stream));
    if (P) poisson_destroy(*P);
    else P = g->pois = new PoissonCtx();
    hipError_t e = poisson_create(*P, g->h.nx, g->h.ny, g->h.nz, dx, dy, dz, g->ctx->stream,
                                  g->ctx->cus);
    if (e != hipSuccess) {
        poisson_destroy(*P);
        delete P;
        g->pois = nullptr;
        return set_err(BSN_ERR_HIP, "Poisson solver set-up (hipFFT): %s", hipGetErrorString(e));
    }
    return BSN_OK;
}

static int ensure_phi(bsn_grid* g) {
    if (!g->d_phi)
        HIPCHK(hipMalloc((void**)&g->d_phi, (size_t)g->h.nx * g->h.ny * g->h.nz * 8));
    return BSN_OK;
}

extern "C" int bsn_solve_poisson_isolated(bsn_grid* g, const double* rho, int where_rho,
                                          double* phi, int where_phi, double dx, double dy,
                                          double dz) {
    if (!g || !rho || !phi) return set_err(BSN_ERR_ARG, "NULL argument");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return set_err(BSN_ERR_ARG, "cell sizes must be > 0");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(ensure_poisson(g, dx, dy, dz));
    const size_t n = (size_t)g->h.nx * g->h.ny * g->h.nz;
    // staging: density in, potential out (host arrays only)
    if (!g->d_rho_in) HIPCHK(hipMalloc((void**)&g->d_rho_in, 2 * n * 8));
    const double* drho = where_rho == BSN_HOST ? g->d_rho_in : rho;
    double* dphi = where_phi == BSN_HOST ? g->d_rho_in + n : phi;
    if (where_rho == BSN_HOST)
        HIPCHK(hipMemcpyAsync(g->d_rho_in, rho, n * 8, hipMemcpyHostToDevice, g->ctx->stream));
    HIPCHK(poisson_solve_array(*g->pois, drho, dphi, g->ctx->stream));
    if (where_phi == BSN_HOST)
        HIPCHK(hipMemcpyAsync(phi, dphi, n * 8, hipMemcpyDeviceToHost, g->ctx->stream));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_grid_set_self_gravity(bsn_grid* g, int on) {
    if (!g || (on != 0 && on != 1)) return set_err(BSN_ERR_ARG, "bad arguments");
    if (on && !g->params_set) return set_err(BSN_ERR_STATE, "grid parameters not set");
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(hipStreamSynchronize(g->ctx->stream));
    if (on) {
        CHK(ensure_poisson(g, g->dx, g->dy, g->dz));
        CHK(ensure_phi(g));
        g->h.phi = g->d_phi;
    } else if (g->self_grav) {
        g->h.phi = nullptr;
    }
    g->self_grav = on;
    return grid_changed(g);
}

extern "C" int bsn_grid_set_phi(bsn_grid* g, const double* phi, int where) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    // a self-gravitating grid solves its own phi every stage into d_phi: a
    // given phi would be overwritten, and freeing d_phi would leave the
    // solve without a destination
    if (g->self_grav)
        return set_err(BSN_ERR_STATE, "self-gravity is on: phi is solved by the step "
                                      "(bsn_grid_set_self_gravity(g, 0) first)");
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(hipStreamSynchronize(g->ctx->stream));
    if (!phi) {
        if (g->d_phi) HIPCHK(hipFree(g->d_phi));
        g->d_phi = nullptr;
        g->h.phi = nullptr;
        return grid_changed(g);
    }
    const size_t bytes = (size_t)g->h.nx * g->h.ny * g->h.nz * 8;
    if (!g->d_phi) HIPCHK(hipMalloc((void**)&g->d_phi, bytes));
    HIPCHK(hipMemcpy(g->d_phi, phi, bytes, where == BSN_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice));
    g->h.phi = g->d_phi;
    return grid_changed(g);
}

static int apply_params(bsn_grid* g, double dx, double dy, double dz, double gamma, int bc,
                        double rho_floor, double P_floor) {
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return set_err(BSN_ERR_ARG, "cell sizes must be > 0");
    if (!(gamma > 1.0)) return set_err(BSN_ERR_ARG, "gamma must be > 1");
    if (bc < 0 || bc > 7) return set_err(BSN_ERR_ARG, "bc mask %d", bc);
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? g->h.nx : a == 1 ? g->h.ny : g->h.nz;
        if ((bc >> a & 1) && n < BSN_NG)
            return set_err(BSN_ERR_ARG, "periodic axis %d needs >= %d cells", a, BSN_NG);
    }
    GridDev& h = g->h;
    const double idx = 1.0 / dx, idy = 1.0 / dy, idz = 1.0 / dz;
    const double gfac = (gamma + 1.0) / (2.0 * gamma);
    const bool same = g->params_set && h.idx == idx && h.idy == idy && h.idz == idz &&
                      h.gamma == gamma && h.bc == bc && h.rho_floor == rho_floor &&
                      h.P_floor == P_floor;
    g->dx = dx; g->dy = dy; g->dz = dz;
    h.idx = idx; h.idy = idy; h.idz = idz;
    h.gamma = gamma; h.gfac = gfac;
    h.gf32[0] = (float)gamma;
    h.gf32[1] = (float)gfac;
    h.gf32[2] = (float)(1.0 / (gamma - 1.0));
    h.bc = bc;
    h.rho_floor = rho_floor;
    h.P_floor = P_floor;
    g->params_set = true;
    if (!same) {
        CHK(grid_changed(g));
        if (g->hs.nbh > 0 || g->hs.nsink > 0) CHK(src_upload(g));
        if (g->self_grav) CHK(ensure_poisson(g, dx, dy, dz));
    }
    return BSN_OK;
}

extern "C" int bsn_grid_set_params(bsn_grid* g, double dx, double dy, double dz, double gamma,
                                   int bc, double rho_floor, double P_floor) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    HIPCHK(hipSetDevice(g->ctx->device));
    return apply_params(g, dx, dy, dz, gamma, bc, rho_floor, P_floor);
}

extern "C" int bsn_grid_device_state(bsn_grid* g, double** U, int64_t* fs, int64_t* pitch) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    if (U) *U = g->h.U;
    if (fs) *fs = g->h.fs;
    if (pitch) *pitch = g->h.pitch;
    return BSN_OK;
}

// ---------------------------------------------------------------------------
// Uniform-grid drop-ins

extern "C" int bsn_fill_ghost_3d(bsn_grid* g, int bc) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    if (bc < 0 || bc > 7) return set_err(BSN_ERR_ARG, "bc mask %d", bc);
    if (g->h.bc != bc) { g->h.bc = bc; CHK(grid_changed(g)); }
    CHK(plan_ghost_fill(*g->plan, 0));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_apply_floors_3d(bsn_grid* g, double rho_floor, double P_floor, double gamma) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    CHK(apply_params(g, g->dx, g->dy, g->dz, gamma, g->h.bc, rho_floor, P_floor));
    CHK(plan_floors(*g->plan, 0));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_euler3d_rhs(bsn_grid* g, double* dU, double* Fx, double* Fy, double* Fz,
                               int where, double dx, double dy, double dz, double gamma, int bc,
                               double rho_floor, double P_floor, int fill_ghosts) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    CHK(apply_params(g, dx, dy, dz, gamma, bc, rho_floor, P_floor));
    Plan& P = *g->plan;
    const LaunchCtx L = g->ctx->L();
    if (fill_ghosts) CHK(plan_ghost_fill(P, 0));
    CHK(plan_floors(P, 0));
    CHK(plan_fluxes(P));
    double* F[3] = {Fx, Fy, Fz};
    for (int d = 0; d < 3; ++d) {
        if (!F[d]) continue;
        const GridDev& h = g->h;
        const size_t n = 5ULL * (h.nx + 2 * BSN_NG + (d == 0)) * (h.Ny + (d == 1)) * (h.Nz + (d == 2));
        if (where == BSN_HOST) {
            CHK(ensure_stage(g, n * 8));
            HIPCHK(launch_flux_export(L, h, d, g->d_stage));
            HIPCHK(hipMemcpyAsync(F[d], g->d_stage, n * 8, hipMemcpyDeviceToHost, g->ctx->stream));
            HIPCHK(hipStreamSynchronize(g->ctx->stream));
        } else {
            HIPCHK(launch_flux_export(L, h, d, F[d]));
        }
    }
    if (dU) {
        CHK(ensure_scratch(g));
        HIPCHK(launch_fill_u64(L, (uint64_t*)g->d_scratch, 5 * g->h.fs, 0));
        CHK(plan_update(P, 0, 0, g->d_scratch));
        CHK(export_ref(g, g->d_scratch, dU, where));
    }
    return sync_ctx(g->ctx);
}

extern "C" int bsn_euler3d_step(bsn_grid* g, double dx, double dy, double dz, double dt,
                                double gamma, int bc, double rho_floor, double P_floor) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    CHK(apply_params(g, dx, dy, dz, gamma, bc, rho_floor, P_floor));
    HIPCHK(launch_set_scal(g->ctx->L(), g->h.scal, SCAL_DT, dt));
    CHK(plan_step(*g->plan, 0));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_cfl_dt_3d(bsn_grid* g, double dx, double dy, double dz, double gamma,
                             double cfl, double* dt_out) {
    if (!g || !dt_out) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    CHK(apply_params(g, dx, dy, dz, gamma, g->h.bc, g->h.rho_floor, g->h.P_floor));
    Plan& P = *g->plan;
    CHK(plan_cfl(P, cfl, 0.0));
    HIPCHK(launch_dt(g->ctx->L(), P.d_grids, 1));
    HIPCHK(hipMemcpyAsync(dt_out, &g->h.scal[SCAL_DT], 8, hipMemcpyDeviceToHost, g->ctx->stream));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_apply_sg_force(bsn_grid* g, double* dU, int where, double dx, double dy,
                                  double dz) {
    if (!g || !dU) return set_err(BSN_ERR_ARG, "NULL argument");
    if (!g->h.phi) return set_err(BSN_ERR_STATE, "grid has no phi");
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(grid_plan(g));
    GridDev h = g->h;
    h.idx = 1.0 / dx; h.idy = 1.0 / dy; h.idz = 1.0 / dz;
    CHK(ensure_scratch(g));
    CHK(import_ref(g, dU, where, g->d_scratch));
    HIPCHK(launch_sg_force(g->ctx->L(), h, g->d_scratch));
    CHK(export_ref(g, g->d_scratch, dU, where));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_rk3_save(bsn_grid* g) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    HIPCHK(hipSetDevice(g->ctx->device));
    HIPCHK(launch_copy_state(g->ctx->L(), g->h, g->h.U, g->h.Un));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_rk3_combine(bsn_grid* g, int stage, const double* dU, int where, double dt) {
    if (!g || !dU) return set_err(BSN_ERR_ARG, "NULL argument");
    if (stage < 1 || stage > 3) return set_err(BSN_ERR_ARG, "stage %d", stage);
    HIPCHK(hipSetDevice(g->ctx->device));
    CHK(ensure_scratch(g));
    CHK(import_ref(g, dU, where, g->d_scratch));
    HIPCHK(launch_combine_all(g->ctx->L(), g->h, stage, g->d_scratch, dt));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_conserved_totals(bsn_grid* g, double dV, double out[5]) {
    if (!g || !out) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    const int nb = TOTALS_BLOCKS;
    if (!g->d_part) HIPCHK(hipMalloc((void**)&g->d_part, nb * 5 * 8));
    HIPCHK(launch_totals(g->ctx->L(), g->h, g->d_part, nb, dV, &g->h.scal[SCAL_TOT]));
    HIPCHK(hipMemcpyAsync(out, &g->h.scal[SCAL_TOT], 5 * 8, hipMemcpyDeviceToHost, g->ctx->stream));
    return sync_ctx(g->ctx);
}

// The same sums with the partials in the fixed scratch of the grid's device
// (set up by bsn_ctx_create): no device memory is allocated.  Every partial
// slot the launch uses is written by its block before the final sum reads
// it, so nothing needs clearing between calls.
extern "C" int bsn_conserved_totals_noalloc(bsn_grid* g, double dV, double* out, int where) {
    if (!g || !out) return set_err(BSN_ERR_ARG, "NULL argument");
    if (where != BSN_HOST && where != BSN_DEVICE) return set_err(BSN_ERR_ARG, "where %d", where);
    double* scr = g->ctx->totals_scratch;
    if (!scr) return set_err(BSN_ERR_STATE, "no totals scratch on device %d", g->ctx->device);
    // the launch uses TOTALS_BLOCKS x 5 partials and 5 sums: within the scratch
    static_assert(5 * TOTALS_BLOCKS + 5 <= TOTALS_SCRATCH_DOUBLES, "totals scratch too small");
    HIPCHK(hipSetDevice(g->ctx->device));
    double* sums = where == BSN_DEVICE ? out : scr + 5 * TOTALS_BLOCKS;
    HIPCHK(launch_totals(g->ctx->L(), g->h, scr, TOTALS_BLOCKS, dV, sums));
    if (where == BSN_DEVICE) return BSN_OK;
    HIPCHK(hipMemcpyAsync(out, sums, 5 * 8, hipMemcpyDeviceToHost, g->ctx->stream));
    return sync_ctx(g->ctx);
}

// ---------------------------------------------------------------------------
// What the bodies and the diagnostics read from the gas (bsn_bodies.hip)

#define BODY_OUT BQ_PART_DOUBLES   // the sums follow the partials

static int body_begin(bsn_grid* g, BodyQ& Q, double dx, double dy, double dz, double x0,
                      double y0, double z0, int n, const double* pos) {
    if (!g) return set_err(BSN_ERR_ARG, "NULL grid");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return set_err(BSN_ERR_ARG, "cell sizes must be > 0");
    if (n < 0 || n > BSN_MAX_BH) return set_err(BSN_ERR_ARG, "body count %d (max %d)", n, BSN_MAX_BH);
    if (n > 0 && !pos) return set_err(BSN_ERR_ARG, "NULL position array");
    Q = BodyQ{};
    Q.n = n;
    Q.x0 = x0; Q.y0 = y0; Q.z0 = z0;
    Q.dx = dx; Q.dy = dy; Q.dz = dz;
    Q.dV = dx * dy * dz;
    for (int b = 0; b < n; ++b) {
        Q.px[b] = pos[3 * b]; Q.py[b] = pos[3 * b + 1]; Q.pz[b] = pos[3 * b + 2];
    }
    HIPCHK(hipSetDevice(g->ctx->device));
    if (!g->d_body) HIPCHK(hipMalloc((void**)&g->d_body, (BODY_OUT + 64) * sizeof(double)));
    return BSN_OK;
}

static int body_finish(bsn_grid* g, hipError_t e, double* host, int count) {
    if (e != hipSuccess) return set_err(BSN_ERR_HIP, "launch: %s", hipGetErrorString(e));
    if (count > 0)
        HIPCHK(hipMemcpyAsync(host, g->d_body + BODY_OUT, count * sizeof(double),
                              hipMemcpyDeviceToHost, g->ctx->stream));
    return sync_ctx(g->ctx);
}

extern "C" int bsn_gas_force_on_bh(bsn_grid* g, double dx, double dy, double dz, int n,
                                   const double* pos, const double* mass, const double* eps,
                                   double x0, double y0, double z0, double* F) {
    BodyQ Q;
    CHK(body_begin(g, Q, dx, dy, dz, x0, y0, z0, n, pos));
    if (n > 0 && (!mass || !eps || !F)) return set_err(BSN_ERR_ARG, "NULL argument");
    if (n == 0) return BSN_OK;
    for (int b = 0; b < n; ++b) { Q.mass[b] = mass[b]; Q.eps[b] = eps[b]; }
    hipError_t e = launch_body_force(g->ctx->stream, g->ctx->cus, g->h, Q, g->d_body,
                                     g->d_body + BODY_OUT);
    return body_finish(g, e, F, 3 * n);
}

// Active cells (0-based first, extent) along one axis whose centres
// o + (ii - 0.5) d can lie within rs of p, one cell of margin each side.  A
// NaN or infinite position or radius makes the reference's `r >= rs` test
// false for every cell or leaves it undecided here, so the whole axis is kept;
// rs <= 0 excludes every cell (r >= 0 >= rs).
static void sink_axis(double p, double rs, double o, double d, int nact, int& lo, int& ext) {
    if (!std::isfinite(p) || !std::isfinite(rs)) { lo = 0; ext = nact; return; }
    if (!(rs > 0.0)) { lo = 0; ext = 0; return; }
    double a = std::floor((p - rs - o) / d + 0.5) - 1.0;      // 1-based
    double b = std::ceil((p + rs - o) / d + 0.5) + 1.0;
    if (!(a >= 1.0)) a = 1.0;
    if (!(b <= (double)nact)) b = (double)nact;
    if (b < a) { lo = 0; ext = 0; return; }
    lo = (int)a - 1;
    ext = (int)b - (int)a + 1;
}

extern "C" int bsn_accrete(bsn_grid* g, double dx, double dy, double dz, int n, const double* pos,
                           double* vel, double* mass, const double* r_sink, const double* t_sink,
                           double rho_sink_min, double x0, double y0, double z0, double dt,
                           double* dm_out) {
    BodyQ Q;
    CHK(body_begin(g, Q, dx, dy, dz, x0, y0, z0, n, pos));
    if (n > 0 && (!vel || !mass || !r_sink || !t_sink)) return set_err(BSN_ERR_ARG, "NULL argument");
    if (n == 0) return BSN_OK;
    Q.rho_sink_min = rho_sink_min;
    const int nact[3] = {g->h.nx, g->h.ny, g->h.nz};
    const double o[3] = {x0, y0, z0}, dd[3] = {dx, dy, dz};
    for (int b = 0; b < n; ++b) {
        Q.rs[b] = r_sink[b];
        Q.rate[b] = Q.dV / t_sink[b] * dt;       // rate = dV / ts * dt (src/sinks.jl:97)
        const double p[3] = {Q.px[b], Q.py[b], Q.pz[b]};
        for (int a = 0; a < 3; ++a) sink_axis(p[a], r_sink[b], o[a], dd[a], nact[a], Q.lo[b][a], Q.ext[b][a]);
    }
    double s[4 * BSN_MAX_BH];
    hipError_t e = launch_body_accrete(g->ctx->stream, g->ctx->cus, g->h, Q, g->d_body,
                                       g->d_body + BODY_OUT);
    CHK(body_finish(g, e, s, 4 * n));
    // the body update of accrete! (src/sinks.jl:104-108)
    for (int b = 0; b < n; ++b) {
        const double dm = s[4 * b];
        if (dm_out) dm_out[b] = dm;
        if (dm > 0.0) {
            const double M_new = mass[b] + dm;
            for (int c = 0; c < 3; ++c) vel[3 * b + c] = (mass[b] * vel[3 * b + c] + s[4 * b + 1 + c]) / M_new;
            mass[b] = M_new;
        }
    }
    return BSN_OK;
}

extern "C" int bsn_gas_diagnostics(bsn_grid* g, double dx, double dy, double dz, double x0,
                                   double y0, double z0, int n, const double* pos,
                                   const double* mass, const double* eps, double gamma,
                                   double out[9]) {
    BodyQ Q;
    CHK(body_begin(g, Q, dx, dy, dz, x0, y0, z0, n, pos));
    if (!out || (n > 0 && (!mass || !eps))) return set_err(BSN_ERR_ARG, "NULL argument");
    for (int b = 0; b < n; ++b) { Q.mass[b] = mass[b]; Q.eps[b] = eps[b]; }
    Q.gamma = gamma;
    double s[BQ_NACC_DIAG];
    hipError_t e = launch_body_diag(g->ctx->stream, g->ctx->cus, g->h, Q, g->d_body,
                                    g->d_body + BODY_OUT);
    CHK(body_finish(g, e, s, BQ_NACC_DIAG));
    for (int q = 0; q < BQ_NACC_DIAG; ++q) out[q] = s[q] * Q.dV;   // Float64(sum(...)) * dV
    return BSN_OK;
}

// ---------------------------------------------------------------------------
// Streaming the state to and from files in k-slabs (bsn_slabs.hip)

extern "C" int bsn_grid_dims(bsn_grid* g, int64_t n[3]) {
    if (!g || !n) return set_err(BSN_ERR_ARG, "NULL argument");
    n[0] = g->h.nx; n[1] = g->h.ny; n[2] = g->h.nz;
    return BSN_OK;
}

struct SlabPlan {
    int parts;
    int64_t K;          // k-planes per part
    int64_t per_k;      // values per k-plane
    int64_t sk;         // k-planes per slab
    int64_t nslab;      // slabs per part
};

static int slab_plan(bsn_grid* g, int layout, int64_t slab_k, SlabPlan& S) {
    const GridDev& h = g->h;
    if (layout == BSN_LAYOUT_FIELDS) {
        S.parts = 5; S.K = h.nz; S.per_k = (int64_t)h.nx * h.ny;
    } else if (layout == BSN_LAYOUT_REF) {
        S.parts = 1; S.K = h.Nz; S.per_k = (int64_t)(h.nx + 2 * BSN_NG) * h.Ny * 5;
    } else {
        return set_err(BSN_ERR_ARG, "unknown layout %d", layout);
    }
    const int64_t target = (32LL << 20) / 8;            // about 32 MB per slab
    int64_t sk = slab_k > 0 ? slab_k : target / S.per_k;
    if (sk < 1) sk = 1;
    if (sk > S.K) sk = S.K;
    while (sk > 1 && sk * S.per_k >= (1LL << 31)) sk /= 2;   // 32-bit element index
    if (sk * S.per_k >= (1LL << 31)) return set_err(BSN_ERR_ARG, "k-plane too large to stream");
    S.sk = sk;
    S.nslab = (S.K + sk - 1) / sk;
    const size_t bytes = (size_t)(sk * S.per_k) * 8;
    CHK(ensure_stage(g, 2 * bytes));
    if (g->pin_bytes < bytes) {
        if (g->h_pin) HIPCHK(hipHostFree(g->h_pin));
        g->h_pin = nullptr;
        g->pin_bytes = 0;
        HIPCHK(hipHostMalloc((void**)&g->h_pin, 2 * bytes, hipHostMallocDefault));
        g->pin_bytes = bytes;
    }
    return BSN_OK;
}

namespace {
struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() { for (auto x : e) if (x) (void)hipEventDestroy(x); }
};
}  // namespace

extern "C" int bsn_grid_export_slabs(bsn_grid* g, int layout, int64_t slab_k, bsn_slab_fn sink,
                                     void* user) {
    if (!g || !sink) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    SlabPlan S;
    CHK(slab_plan(g, layout, slab_k, S));
    hipStream_t st = g->ctx->stream;
    const int64_t slab_vals = S.sk * S.per_k;
    double* pin[2] = {g->h_pin, g->h_pin + slab_vals};
    double* dev[2] = {g->d_stage, g->d_stage + slab_vals};
    Events ev;
    for (auto& x : ev.e) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    const int64_t total = S.parts * S.nslab;
    int prev_part = 0;
    int64_t prev_k0 = 0, prev_nk = 0;
    for (int64_t s = 0; s <= total; ++s) {
        const int b = (int)(s & 1);
        int part = 0;
        int64_t k0 = 0, nk = 0;
        if (s < total) {
            part = (int)(s / S.nslab);
            k0 = (s % S.nslab) * S.sk;
            nk = std::min(S.sk, S.K - k0);
            HIPCHK(launch_slab(st, g->ctx->cus, g->h, 1, layout, part, (int)k0, (int)nk, dev[b]));
            HIPCHK(hipMemcpyAsync(pin[b], dev[b], (size_t)(nk * S.per_k) * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(hipEventRecord(ev.e[b], st));
        }
        if (s > 0) {
            // the previous slab is on the host once its copy has finished;
            // the GPU meanwhile packs and copies slab s
            HIPCHK(hipEventSynchronize(ev.e[1 - b]));
            if (sink(user, prev_part, prev_k0, prev_nk, pin[1 - b], prev_nk * S.per_k) != 0) {
                (void)hipStreamSynchronize(st);
                return set_err(BSN_ERR_IO, "slab consumer failed (part %d, k0 %lld)", prev_part,
                               (long long)prev_k0);
            }
        }
        prev_part = part; prev_k0 = k0; prev_nk = nk;
    }
    return sync_ctx(g->ctx);
}

extern "C" int bsn_grid_import_slabs(bsn_grid* g, int layout, int64_t slab_k, bsn_slab_fn source,
                                     void* user) {
    if (!g || !source) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(g->ctx->device));
    SlabPlan S;
    CHK(slab_plan(g, layout, slab_k, S));
    hipStream_t st = g->ctx->stream;
    const int64_t slab_vals = S.sk * S.per_k;
    double* pin[2] = {g->h_pin, g->h_pin + slab_vals};
    double* dev[2] = {g->d_stage, g->d_stage + slab_vals};
    Events ev;
    for (auto& x : ev.e) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    const int64_t total = S.parts * S.nslab;
    for (int64_t s = 0; s < total; ++s) {
        const int b = (int)(s & 1);
        const int part = (int)(s / S.nslab);
        const int64_t k0 = (s % S.nslab) * S.sk;
        const int64_t nk = std::min(S.sk, S.K - k0);
        // buffer b is free once slab s - 2 has been copied up and unpacked
        if (s >= 2) HIPCHK(hipEventSynchronize(ev.e[b]));
        if (source(user, part, k0, nk, pin[b], nk * S.per_k) != 0) {
            (void)hipStreamSynchronize(st);
            return set_err(BSN_ERR_IO, "slab producer failed (part %d, k0 %lld)", part, (long long)k0);
        }
        HIPCHK(hipMemcpyAsync(dev[b], pin[b], (size_t)(nk * S.per_k) * 8, hipMemcpyHostToDevice, st));
        HIPCHK(launch_slab(st, g->ctx->cus, g->h, 0, layout, part, (int)k0, (int)nk, dev[b]));
        HIPCHK(hipEventRecord(ev.e[b], st));
    }
    return sync_ctx(g->ctx);
}

// Host-array drop-ins with a per-thread cached context and grid.
static thread_local bsn_ctx* t_ctx = nullptr;
static thread_local bsn_grid* t_grid = nullptr;

static int host_grid(int64_t nx, int64_t ny, int64_t nz, int prec) {
    if (!t_ctx) CHK(bsn_ctx_create(0, &t_ctx));
    if (t_grid && (t_grid->h.nx != nx || t_grid->h.ny != ny || t_grid->h.nz != nz || t_grid->h.prec != prec)) {
        bsn_grid_destroy(t_grid);
        t_grid = nullptr;
    }
    if (!t_grid) CHK(bsn_grid_create(t_ctx, nx, ny, nz, prec, &t_grid));
    return BSN_OK;
}

extern "C" int bsn_euler3d_step_host(double* U, int64_t nx, int64_t ny, int64_t nz, double dx,
                                     double dy, double dz, double dt, double gamma, int bc,
                                     double rho_floor, double P_floor, int precision) {
    CHK(host_grid(nx, ny, nz, precision));
    CHK(bsn_grid_set_state(t_grid, U, BSN_HOST));
    CHK(bsn_euler3d_step(t_grid, dx, dy, dz, dt, gamma, bc, rho_floor, P_floor));
    return bsn_grid_get_state(t_grid, U, BSN_HOST);
}

extern "C" int bsn_cfl_dt_3d_host(const double* U, int64_t nx, int64_t ny, int64_t nz, double dx,
                                  double dy, double dz, double gamma, double cfl, double* dt_out) {
    CHK(host_grid(nx, ny, nz, t_grid ? t_grid->h.prec : 0));
    CHK(bsn_grid_set_state(t_grid, U, BSN_HOST));
    return bsn_cfl_dt_3d(t_grid, dx, dy, dz, gamma, cfl, dt_out);
}

// ---------------------------------------------------------------------------
// Batches

extern "C" int bsn_batch_create(bsn_ctx* ctx, int ngrids, bsn_grid* const* grids, bsn_batch** out) {
    if (!ctx || !grids || !out || ngrids < 1) return set_err(BSN_ERR_ARG, "bad arguments");
    std::vector<bsn_grid*> v(grids, grids + ngrids);
    for (auto* g : v)
        if (!g || !g->params_set) return set_err(BSN_ERR_STATE, "grid without parameters");
    HIPCHK(hipSetDevice(ctx->device));
    bsn_batch* b = new bsn_batch();
    int r = plan_new(ctx, v, &b->plan);
    if (r != BSN_OK) { delete b; return r; }
    *out = b;
    return BSN_OK;
}

extern "C" int bsn_batch_destroy(bsn_batch* b) {
    if (!b) return BSN_OK;
    plan_free(b->plan);
    delete b;
    return BSN_OK;
}

extern "C" int bsn_batch_set_shapes(bsn_batch* b, int sx, int sy, int sz) {
    if (!b) return set_err(BSN_ERR_ARG, "NULL batch");
    const int s[3] = {sx, sy, sz};
    for (int d = 0; d < 3; ++d)
        if (s[d] < 0 || s[d] >= flux_shape_count(d)) return set_err(BSN_ERR_ARG, "shape %d", s[d]);
    HIPCHK(hipSetDevice(b->plan->ctx->device));
    for (int d = 0; d < 3; ++d) CHK(plan_set_shape(*b->plan, d, s[d]));
    return BSN_OK;
}

extern "C" int bsn_batch_set_update_shape(bsn_batch* b, int shape) {
    if (!b) return set_err(BSN_ERR_ARG, "NULL batch");
    if (shape < 0 || shape >= zupdate_shape_count()) return set_err(BSN_ERR_ARG, "shape %d", shape);
    HIPCHK(hipSetDevice(b->plan->ctx->device));
    return plan_set_zshape(*b->plan, shape);
}

extern "C" int bsn_batch_get_shapes(bsn_batch* b, int shapes[4]) {
    if (!b || !shapes) return set_err(BSN_ERR_ARG, "NULL argument");
    for (int d = 0; d < 3; ++d) shapes[d] = b->plan->shape[d];
    shapes[3] = b->plan->zshape;
    return BSN_OK;
}

extern "C" int bsn_batch_set_mode(bsn_batch* b, int inplace) {
    if (!b || (inplace != 0 && inplace != 1)) return set_err(BSN_ERR_ARG, "bad arguments");
    b->plan->inplace = inplace;
    graph_drop(b->plan->gexec, b->plan->graph);
    return BSN_OK;
}

// Times every tile shape of every kernel of a plan on its current states and
// keeps the fastest; appends a JSON object of the timings to rep.
// Name of a flux-sweep shape in the autotuner's report: "WxH" for a tiled
// shape, "walkL" for a register walk of L interfaces.
static std::string flux_shape_name(int dir, int shape) {
    if (flux_walk_len(dir, shape) > 0)
        return "walk" + std::to_string(flux_walk_len(dir, shape)) +
               (dir == 0 && shape == FLUX_XWALK_SHAPE0 + 2 ? "d" : "");
    int W, H;
    flux_shape_dims(dir, shape, W, H);
    return std::to_string(W) + "x" + std::to_string(H);
}

// Name of a fused-kernel shape in the autotuner's report: segmentation
// (kz16, kz32, kbal) and variant.
static const char* zshape_name(int prec, int shape) {
    static const char* names[2][12] = {
        {"kz16", "kz32", "kbal", "kz16pf", "kz32pf", "kbalpf", "kz16ring", "kz32ring", "kbalring",
         "kz16ringpf", "kz32ringpf", "kbalringpf"},
        {"kz16", "kz32", "kbal", "kz16ringpf2w", "kz32ringpf2w", "kbalringpf2w", "kz16ring",
         "kz32ring", "kbalring", "kz16ringpf", "kz32ringpf", "kbalringpf"}};
    return names[prec == 1][shape];
}

static int plan_autotune(Plan& P, std::string& out) {
    HIPCHK(hipSetDevice(P.ctx->device));
    CHK(plan_refresh(P));
    const LaunchCtx L = P.ctx->L();
    // Every kernel is timed inside the out-of-place step's own sequence (x
    // sweep, y sweep, fused update) on stages 1 and 2: stage 1 reads the
    // caller's state A (ghosts mapped, floors on load) and writes B, stage 2
    // reads B and writes C, so A is never written.  Events bracket the
    // candidate only; its time is (stage 1 + 2 x stage 2) / 3, as stage 3
    // runs like stage 2 (no floors on load, the base state read).  A kernel
    // launched back to back with itself, or on stage 1 alone, sees other
    // caches, tails and work than it does in the step: a tiled x sweep timed
    // alone ran 5 % slower than the x walk and 4 % faster inside the step.
    const GridDev* T1 = P.d_st[0];
    const int reps = 3;
    std::vector<hipEvent_t> ev(4 * reps);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    auto seq = [&](int which, int stage, hipEvent_t* e) -> int {
        // one stage: x sweep, y sweep, fused update; 'which' (0 x, 1 y,
        // 3 fused) is bracketed by e[0], e[1] when e is given
        const GridDev* T = P.d_st[stage - 1];
        for (int k : {0, 1, 3}) {
            if (e && k == which) HIPCHK(hipEventRecord(e[0], L.stream));
            if (k < 3) HIPCHK(launch_flux(L, k, P.prec, P.shape[k], T, P.d_ft[k], P.nft[k], 0));
            else       HIPCHK(launch_zupdate(L, P.prec, P.zshape, T, P.d_zt, P.nzt, stage, 0));
            if (e && k == which) HIPCHK(hipEventRecord(e[1], L.stream));
        }
        return BSN_OK;
    };
    auto time_in_seq = [&](int which, float& ms) -> int {
        CHK(seq(which, 1, nullptr));                    // warm-up passes
        CHK(seq(which, 2, nullptr));
        for (int r = 0; r < reps; ++r) {
            CHK(seq(which, 1, &ev[4 * r]));
            CHK(seq(which, 2, &ev[4 * r + 2]));
        }
        HIPCHK(hipEventSynchronize(ev[4 * reps - 1]));
        float m1 = 0, m2 = 0;
        for (int r = 0; r < reps; ++r) {
            float a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, ev[4 * r], ev[4 * r + 1]));
            HIPCHK(hipEventElapsedTime(&b, ev[4 * r + 2], ev[4 * r + 3]));
            m1 += a;
            m2 += b;
        }
        ms = (m1 + 2.0f * m2) / (3.0f * reps);
        return BSN_OK;
    };
    // The GPU comes to the tuner idle: its clocks and TLBs settle over the
    // first few milliseconds of work, long enough to slow whichever candidate
    // is timed first by a tenth.  Whole stages run first, for at least 20 ms,
    // and the candidates are then timed in turn, keeping each one's best of
    // three passes.
    {
        HIPCHK(hipEventRecord(ev[0], L.stream));
        for (int n = 0;; ++n) {
            CHK(seq(-1, 1 + (n & 1), nullptr));
            if (n % 8 != 7) continue;
            HIPCHK(hipEventRecord(ev[1], L.stream));
            HIPCHK(hipEventSynchronize(ev[1]));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (ms >= 20.0f || n >= 4000) break;
        }
    }
    const int passes = 3;
    std::string rep = "{";
    auto pick = [&](const char* key, int which, int count, auto set, auto name) -> int {
        // time every shape of one kernel in the sequence, in turn; keep the best
        std::vector<float> t(count, 1e30f);
        for (int r = 0; r < passes; ++r)
            for (int s = 0; s < count; ++s) {
                CHK(set(s));
                float ms = 0;
                if (which >= 0) {
                    CHK(time_in_seq(which, ms));
                } else {
                    // the z sweep runs only in the in-place form: timed alone
                    HIPCHK(launch_flux(L, 2, P.prec, s, T1, P.d_ft[2], P.nft[2], 0));
                    HIPCHK(hipEventRecord(ev[0], L.stream));
                    for (int q = 0; q < reps; ++q)
                        HIPCHK(launch_flux(L, 2, P.prec, s, T1, P.d_ft[2], P.nft[2], 0));
                    HIPCHK(hipEventRecord(ev[1], L.stream));
                    HIPCHK(hipEventSynchronize(ev[1]));
                    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
                    ms /= reps;
                }
                t[s] = std::min(t[s], ms);
            }
        int best = 0;
        char buf[160];
        rep += std::string(rep.size() > 1 ? ", " : "") + "\"" + key + "\": {";
        for (int s = 0; s < count; ++s) {
            snprintf(buf, sizeof buf, "%s\"%s\": %.4f", s ? ", " : "", name(s).c_str(), t[s]);
            rep += buf;
            if (t[s] < t[best]) best = s;
        }
        CHK(set(best));
        snprintf(buf, sizeof buf, ", \"best\": \"%s\"}", name(best).c_str());
        rep += buf;
        return BSN_OK;
    };
    for (int d = 0; d < 3; ++d)
        CHK(pick(d == 0 ? "x" : d == 1 ? "y" : "z", d < 2 ? d : -1, flux_shape_count(d),
                 [&](int s) { return plan_set_shape(P, d, s); },
                 [&](int s) { return flux_shape_name(d, s); }));
    // fused z sweep + update, in the sequence with the kept x and y shapes
    CHK(pick("zupdate", 3, zupdate_shape_count(),
             [&](int s) { return plan_set_zshape(P, s); },
             [&](int s) { return std::string(zshape_name(P.prec, s)); }));
    rep += "}";
    HIPCHK(hipStreamSynchronize(L.stream));
    for (auto e : ev) (void)hipEventDestroy(e);
    out += rep;
    return BSN_OK;
}

extern "C" int bsn_batch_autotune(bsn_batch* b, char* report, int report_len) {
    if (!b) return set_err(BSN_ERR_ARG, "NULL batch");
    std::string rep;
    CHK(plan_autotune(*b->plan, rep));
    if (report && report_len > 0) snprintf(report, (size_t)report_len, "%s", rep.c_str());
    return BSN_OK;
}

extern "C" int bsn_batch_step(bsn_batch* b, const double* dt) {
    if (!b || !dt) return set_err(BSN_ERR_ARG, "NULL argument");
    Plan& P = *b->plan;
    HIPCHK(hipSetDevice(P.ctx->device));
    CHK(plan_refresh(P));
    for (size_t g = 0; g < P.grids.size(); ++g)
        HIPCHK(launch_set_scal(P.ctx->L(), P.grids[g]->h.scal, SCAL_DT, dt[g]));
    CHK(plan_step(P, 0));
    return sync_ctx(P.ctx);
}

// One device-dt step: dt from the running maximum, RK3 with the stage-3
// update accumulating the next maximum, then t += dt.
static int plan_device_step(Plan& P) {
    const LaunchCtx L = P.ctx->L();
    HIPCHK(launch_dt(L, P.d_grids, (int)P.grids.size()));
    CHK(plan_step(P, 1));
    HIPCHK(launch_advance_t(L, P.d_grids, (int)P.grids.size()));
    return BSN_OK;
}

static int plan_run(Plan& P, double cfl, double t_end, int nsteps, int use_graph) {
    const LaunchCtx L = P.ctx->L();
    for (auto* g : P.grids) HIPCHK(launch_set_scal(L, g->h.scal, SCAL_T, 0.0));
    CHK(plan_cfl(P, cfl, t_end));
    bool fft = false;
    for (auto* g : P.grids) fft = fft || g->self_grav;
    if (!use_graph || fft) {     // FFT plans are launched eagerly
        for (int s = 0; s < nsteps; ++s) CHK(plan_device_step(P));
        return sync_ctx(P.ctx);
    }
    if (!P.gexec) {
        HIPCHK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
        int r = plan_device_step(P);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(L.stream, &gr);
        if (r != BSN_OK) { if (gr) (void)hipGraphDestroy(gr); return r; }
        if (e != hipSuccess) return set_err(BSN_ERR_HIP, "capture: %s", hipGetErrorString(e));
        P.graph = gr;
        HIPCHK(hipGraphInstantiate(&P.gexec, P.graph, nullptr, nullptr, 0));
    }
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(P.gexec, L.stream));
    return sync_ctx(P.ctx);
}

extern "C" int bsn_batch_run(bsn_batch* b, double cfl, double t_end, int nsteps, int use_graph) {
    if (!b || nsteps < 0) return set_err(BSN_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(b->plan->ctx->device));
    CHK(plan_refresh(*b->plan));
    return plan_run(*b->plan, cfl, t_end, nsteps, use_graph);
}

extern "C" int bsn_batch_times(bsn_batch* b, double* t, double* dt) {
    if (!b) return set_err(BSN_ERR_ARG, "NULL batch");
    Plan& P = *b->plan;
    HIPCHK(hipSetDevice(P.ctx->device));
    for (size_t g = 0; g < P.grids.size(); ++g) {
        double s[SCAL_COUNT];
        HIPCHK(hipMemcpy(s, P.grids[g]->h.scal, sizeof s, hipMemcpyDeviceToHost));
        if (t) t[g] = s[SCAL_T];
        if (dt) dt[g] = s[SCAL_DT];
    }
    return BSN_OK;
}

extern "C" int bsn_batch_profile(bsn_batch* b, int nsteps, double* ms_flux3, double* ms_update,
                                 double* ms_step) {
    if (!b || nsteps < 1) return set_err(BSN_ERR_ARG, "bad arguments");
    Plan& P = *b->plan;
    HIPCHK(hipSetDevice(P.ctx->device));
    CHK(plan_refresh(P));
    const LaunchCtx L = P.ctx->L();
    hipEvent_t s0, s1;
    HIPCHK(hipEventCreate(&s0));
    HIPCHK(hipEventCreate(&s1));
    P.prof = true;
    HIPCHK(hipEventRecord(s0, L.stream));
    int r = BSN_OK;
    for (int s = 0; s < nsteps && r == BSN_OK; ++s) r = plan_device_step(P);
    P.prof = false;
    HIPCHK(hipEventRecord(s1, L.stream));
    HIPCHK(hipEventSynchronize(s1));
    double acc[4] = {0, 0, 0, 0};
    int cnt[4] = {0, 0, 0, 0};
    for (size_t e = 0; e + 1 < P.ev.size(); ++e) {
        const int k = P.ev_kind[e];
        if (k > 3) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, P.ev[e], P.ev[e + 1]) == hipSuccess) { acc[k] += ms; cnt[k]++; }
    }
    for (auto e : P.ev) (void)hipEventDestroy(e);
    P.ev.clear();
    P.ev_kind.clear();
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, s0, s1));
    (void)hipEventDestroy(s0);
    (void)hipEventDestroy(s1);
    if (r != BSN_OK) return r;
    if (ms_flux3) for (int d = 0; d < 3; ++d) ms_flux3[d] = cnt[d] ? acc[d] / cnt[d] : 0.0;
    if (ms_update) *ms_update = cnt[3] ? acc[3] / cnt[3] : 0.0;
    if (ms_step) *ms_step = tot / nsteps;
    return BSN_OK;
}

// ---------------------------------------------------------------------------
// FMR

// nearest coarse active index of a fine index (_coarse_stencil, fmr3d.jl:165-170)
static int coarse_centre(int ig_f1, int c_lo, int ratio) {
    const double xi = ((double)c_lo - 0.5) + ((double)(ig_f1 - BSN_NG) - 0.5) / (double)ratio;
    return (int)rint(xi);
}

extern "C" int bsn_fmr_create(bsn_ctx* ctx, int nlevels, bsn_grid* const* levels,
                              const int64_t* lo_hi, const int* ratio, bsn_fmr** out) {
    if (!ctx || !levels || !out || nlevels < 1 || nlevels > BSN_MAXLEV)
        return set_err(BSN_ERR_ARG, "bad arguments");
    if (nlevels > 1 && (!lo_hi || !ratio)) return set_err(BSN_ERR_ARG, "missing boxes");
    HIPCHK(hipSetDevice(ctx->device));
    for (int l = 0; l < nlevels; ++l) {
        if (!levels[l] || !levels[l]->params_set) return set_err(BSN_ERR_STATE, "level %d without parameters", l);
        if (levels[l]->h.prec != levels[0]->h.prec) return set_err(BSN_ERR_ARG, "mixed precision");
    }
    for (int l = 0; l + 1 < nlevels; ++l) {
        const GridDev& P = levels[l]->h;
        const GridDev& C = levels[l + 1]->h;
        const int64_t* b = &lo_hi[6 * l];
        const int r = ratio[l];
        if (r < 1) return set_err(BSN_ERR_ARG, "ratio %d", r);
        const int n[3] = {P.nx, P.ny, P.nz}, nc[3] = {C.nx, C.ny, C.nz};
        if (r > 4) return set_err(BSN_ERR_ARG, "ratio %d > 4 not supported", r);
        for (int a = 0; a < 3; ++a) {
            const int64_t lo = b[2 * a], hi = b[2 * a + 1];
            if (lo < 1 || hi > n[a] || hi < lo) return set_err(BSN_ERR_ARG, "level %d box axis %d", l + 1, a);
            if ((hi - lo + 1) * r != nc[a]) return set_err(BSN_ERR_ARG, "level %d size mismatch axis %d", l + 1, a);
            // every 5-point prolongation stencil (fmr3d.jl:165-198) must stay
            // inside the parent's padded array: check the first and last fine
            // index, between which the coarse centre is monotone
            const int first = coarse_centre(1, (int)lo, r);
            const int last = coarse_centre(nc[a] + 2 * BSN_NG, (int)lo, r);
            if (BSN_NG + first - 3 < 0 || BSN_NG + last + 1 > n[a] + 2 * BSN_NG - 1)
                return set_err(BSN_ERR_ARG, "level %d box too close to the parent edge (axis %d)", l + 1, a);
        }
    }
    bsn_fmr* m = new bsn_fmr();
    m->ctx = ctx;
    m->nlev = nlevels;
    m->lev.assign(levels, levels + nlevels);
    m->alpha.assign(nlevels, 0.0);
    // the z interfaces a level's flux registers read, which the fused step
    // stores: a parent's faces at its child's box (interfaces ck_lo - 1 and
    // ck_hi of k_fr_coarse), a child's outer faces (0 and nz of k_fr_fine)
    for (int l = 0; l < nlevels; ++l) {
        GridDev& h = levels[l]->h;
        h.nzkeep = 0;
        if (l + 1 < nlevels) {
            h.zkeep[h.nzkeep++] = (int)lo_hi[6 * l + 4] - 1;
            h.zkeep[h.nzkeep++] = (int)lo_hi[6 * l + 5];
        }
        if (l > 0) {
            h.zkeep[h.nzkeep++] = 0;
            h.zkeep[h.nzkeep++] = h.nz;
        }
    }
    auto fail = [&](int code) {
        bsn_fmr_destroy(m);
        return code;
    };
    for (int l = 0; l < nlevels; ++l) {
        Plan* P = nullptr;
        int r = plan_new(ctx, {levels[l]}, &P);
        if (r != BSN_OK) return fail(r);
        m->plans.push_back(P);
    }
    std::vector<GridDev> all;
    for (auto* g : m->lev) all.push_back(g->h);
    if (upload_vec(&m->d_all, all) != BSN_OK) return fail(BSN_ERR_HIP);
    for (int l = 0; l + 1 < nlevels; ++l) {
        FmrDev f{};
        f.parent = l;
        f.child = l + 1;
        f.ratio = ratio[l];
        for (int a = 0; a < 3; ++a) { f.lo[a] = (int)lo_hi[6 * l + 2 * a]; f.hi[a] = (int)lo_hi[6 * l + 2 * a + 1]; }
        // src/fmr3d.jl:32-46
        const double nodes[5] = {-2.0, -1.0, 0.0, 1.0, 2.0};
        for (int s = 0; s < f.ratio && s < 4; ++s) {
            const double t = ((double)s + 0.5 - (double)f.ratio / 2.0) / (double)f.ratio;
            for (int j = 0; j < 5; ++j) {
                double wj = 1.0;
                for (int mm = 0; mm < 5; ++mm)
                    if (mm != j) wj *= (t - nodes[mm]) / (nodes[j] - nodes[mm]);
                f.w[s * 5 + j] = wj;
            }
        }
        if (f.ratio > 4) return fail(set_err(BSN_ERR_ARG, "ratio > 4 not supported"));
        const int nic = f.hi[0] - f.lo[0] + 1, njc = f.hi[1] - f.lo[1] + 1, nkc = f.hi[2] - f.lo[2] + 1;
        const int64_t face[6] = {(int64_t)njc * nkc, (int64_t)njc * nkc, (int64_t)nic * nkc,
                                 (int64_t)nic * nkc, (int64_t)nic * njc, (int64_t)nic * njc};
        f.froff[0] = 0;
        for (int k = 0; k < 6; ++k) f.froff[k + 1] = f.froff[k] + 5 * face[k];
        if (hipMalloc((void**)&f.frc, f.froff[6] * 8) != hipSuccess ||
            hipMalloc((void**)&f.frf, f.froff[6] * 8) != hipSuccess)
            return fail(set_err(BSN_ERR_HIP, "flux register allocation"));
        f.Ucn = levels[l]->h.Un;
        f.Ucnp1 = levels[l]->h.U;
        // the parent cells any prolongation stencil of the child reads (the
        // same first / last fine index bound as the box check above)
        for (int a = 0; a < 3; ++a) {
            const GridDev& ch = levels[l + 1]->h;
            const int nca = a == 0 ? ch.nx : a == 1 ? ch.ny : ch.nz;
            f.plo[a] = BSN_NG + coarse_centre(1, f.lo[a], f.ratio) - 3;
            f.phi[a] = BSN_NG + coarse_centre(nca + 2 * BSN_NG, f.lo[a], f.ratio) + 1;
        }
        if (hipMalloc((void**)&f.Uci, 5 * levels[l]->h.fs * 8) != hipSuccess)
            return fail(set_err(BSN_ERR_HIP, "interpolation buffer allocation"));
        const int64_t box = (int64_t)(f.phi[0] - f.plo[0] + 1) * (f.phi[1] - f.plo[1] + 1) *
                            (f.phi[2] - f.plo[2] + 1);
        if (hipMalloc((void**)&f.Ucit, 5 * box * 8) != hipSuccess)
            return fail(set_err(BSN_ERR_HIP, "interpolation buffer allocation"));
        m->hf.push_back(f);
    }
    if (!m->hf.empty() && upload_vec(&m->d_f, m->hf) != BSN_OK) return fail(BSN_ERR_HIP);
    *out = m;
    return BSN_OK;
}

extern "C" int bsn_fmr_destroy(bsn_fmr* m) {
    if (!m) return BSN_OK;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    graph_drop(m->gexec, m->graph);
    for (auto* P : m->plans) plan_free(P);
    for (auto& f : m->hf) {
        (void)hipFree(f.frc); (void)hipFree(f.frf); (void)hipFree(f.Uci); (void)hipFree(f.Ucit);
    }
    for (auto* t : m->d_fst) (void)hipFree(t);
    for (auto* p : m->d_sgphi) (void)hipFree(p);
    for (auto* g : m->lev) g->h.nzkeep = 0;
    (void)hipFree(m->d_f);
    (void)hipFree(m->d_all);
    delete m;
    return BSN_OK;
}

// Stage tables of the fused level step.  Level buffers: U (the state), Un
// (its raw copy at the start of the step, the RK3 base state and, for a
// parent, the start state its child interpolates from), W1 and W2 (work):
// stage 1 reads U and writes W1, stage 2 reads W1 and writes W2, stage 3
// reads W2 and writes U, so U ends with the new state and Un keeps the old
// one as in the in-place step.  The root maps ghost cells to active cells
// under its boundary conditions on load; a child reads the ghost cells
// prolonged into each stage's input buffer.  Stage 1 floors active cells on
// load (the start-of-stage floors).
static int fmr_stage_tables(bsn_fmr* m) {
    if (m->d_fst.empty()) m->d_fst.assign(3 * m->nlev, nullptr);
    for (int l = 0; l < m->nlev; ++l) {
        bsn_grid* g = m->lev[l];
        CHK(ensure_work(g));
        CHK(ensure_work2(g));
        const int map = l == 0 ? 1 : 0;
        GridDev d = g->h;
        d.Un = g->h.Un;
        GridDev st[3] = {d, d, d};
        st[0].lmode = map | 2; st[0].U = g->h.U;   st[0].Uout = g->d_work;
        st[1].lmode = map;     st[1].U = g->d_work; st[1].Uout = g->d_work2;
        st[2].lmode = map;     st[2].U = g->d_work2; st[2].Uout = g->h.U;
        for (int s = 0; s < 3; ++s) {
            GridDev*& t = m->d_fst[3 * l + s];
            if (!t) HIPCHK(hipMalloc((void**)&t, sizeof(GridDev)));
            HIPCHK(hipMemcpy(t, &st[s], sizeof(GridDev), hipMemcpyHostToDevice));
        }
    }
    return BSN_OK;
}

static int fmr_refresh(bsn_fmr* m) {
    std::vector<GridDev> all;
    for (auto* g : m->lev) all.push_back(g->h);
    HIPCHK(hipMemcpy(m->d_all, all.data(), all.size() * sizeof(GridDev), hipMemcpyHostToDevice));
    for (auto* P : m->plans) CHK(plan_refresh(*P));
    if (!m->inplace) CHK(fmr_stage_tables(m));
    return BSN_OK;
}

static int64_t ghost_total(const GridDev& d) {
    return 2 * BSN_NG * ((int64_t)d.Ny * d.Nz + (int64_t)d.nx * d.Nz + (int64_t)d.nx * d.ny);
}

// interp: from the parent state interpolated in time (fmr_interp), else from
// the parent's current state.
static int fmr_prolong(bsn_fmr* m, int l, int all_cells, bool interp,
                       ProlongDst dst = ProlongDst{0, {}}) {
    const GridDev& C = m->lev[l]->h;
    const GridDev& Pg = m->lev[l - 1]->h;
    const int64_t total = all_cells ? (int64_t)(C.nx + 2 * BSN_NG) * C.Ny * C.Nz : ghost_total(C);
    HIPCHK(launch_prolong(m->ctx->L(), m->d_all, m->d_f + (l - 1), total,
                          interp ? m->hf[l - 1].Uci : Pg.U, interp ? m->hf[l - 1].Ucit : nullptr,
                          all_cells, dst));
    return BSN_OK;
}

// The parent state of level l at time fraction alpha[l] of its step, over
// the cells the prolongation reads; once per sub-cycle of level l.
static int fmr_interp(bsn_fmr* m, int l) {
    const FmrDev& f = m->hf[l - 1];
    const int64_t total = (int64_t)(f.phi[0] - f.plo[0] + 1) * (f.phi[1] - f.plo[1] + 1) *
                          (f.phi[2] - f.plo[2] + 1);
    HIPCHK(launch_time_interp(m->ctx->L(), m->d_all, m->d_f + (l - 1), total, m->alpha[l]));
    return BSN_OK;
}

static const double RK_W[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};   // fmr3d.jl:496

// Advance level l by its dt (already in its scal): own RK3 with flux-register
// accumulation, then the child's sub-cycles, restriction and correction.
static int fmr_advance(bsn_fmr* m, int l) {
    Plan& P = *m->plans[l];
    const LaunchCtx L = m->ctx->L();
    const bool child = l + 1 < m->nlev;
    const GridDev& h = m->lev[l]->h;
    if (child) HIPCHK(launch_fr_zero(L, m->d_f + l, m->hf[l].froff[6]));
    if (!m->inplace) {
        // fused form: per stage the x and y sweeps, then the z sweep fused
        // with the update (storing the z faces the registers read), then the
        // flux-register sums of the stage; same bits as the in-place form
        HIPCHK(launch_copy_state(L, h, h.U, h.Un));
        if (l > 0) {
            // alpha and the interpolated parent are fixed for the sub-cycle,
            // so the ghost cells of all three stage inputs are prolonged once
            CHK(fmr_interp(m, l));
            bsn_grid* g = m->lev[l];
            CHK(fmr_prolong(m, l, 0, true, ProlongDst{3, {g->h.U, g->d_work, g->d_work2}}));
        }
        for (int s = 1; s <= 3; ++s) {
            const GridDev* T = m->d_fst[3 * l + s - 1];
            CHK(plan_fluxes(P, T, 2));
            prof_mark(P, 3);
            HIPCHK(launch_zupdate(L, P.prec, P.zshape, T, P.d_zt, P.nzt, s, 0));
            prof_mark(P, 9);
            if (child) HIPCHK(launch_fr_coarse(L, P.prec, m->d_all, m->d_f + l, m->hf[l].froff[6] / 5, RK_W[s - 1]));
            if (l > 0) HIPCHK(launch_fr_fine(L, P.prec, m->d_all, m->d_f + (l - 1), m->hf[l - 1].froff[6] / 5, RK_W[s - 1]));
        }
        CHK(plan_ghost_combine(P, 3, m->d_fst[3 * l + 2]));
    } else {
    if (l > 0) HIPCHK(launch_copy_state(L, h, h.U, h.Un));
    if (l > 0) CHK(fmr_interp(m, l));
    for (int s = 1; s <= 3; ++s) {
        if (l == 0) {
            if (s == 1) { CHK(plan_ghost_fill(P, 1)); CHK(plan_floors(P, 1)); }
            else CHK(plan_ghost_fill(P, 0));
        } else {
            CHK(fmr_prolong(m, l, 0, true));
            CHK(plan_floors(P, 0));
        }
        CHK(plan_fluxes(P));
        if (child) HIPCHK(launch_fr_coarse(L, P.prec, m->d_all, m->d_f + l, m->hf[l].froff[6] / 5, RK_W[s - 1]));
        if (l > 0) HIPCHK(launch_fr_fine(L, P.prec, m->d_all, m->d_f + (l - 1), m->hf[l - 1].froff[6] / 5, RK_W[s - 1]));
        CHK(plan_update(P, s, 0));
    }
    CHK(plan_ghost_combine(P, 3));
    }
    if (child) {
        const FmrDev& f = m->hf[l];
        HIPCHK(launch_child_dt(L, m->d_all, l, l + 1, f.ratio));
        for (int ss = 0; ss < f.ratio; ++ss) {
            m->alpha[l + 1] = (double)ss / (double)f.ratio;
            CHK(fmr_advance(m, l + 1));
        }
        const int64_t ncells = (int64_t)(f.hi[0] - f.lo[0] + 1) * (f.hi[1] - f.lo[1] + 1) * (f.hi[2] - f.lo[2] + 1);
        HIPCHK(launch_restrict(L, m->d_all, m->d_f + l, ncells));
        HIPCHK(launch_bc_correct(L, m->d_all, m->d_f + l, f.froff[6] / 5));
        CHK(plan_floors(P, 0));
    }
    return BSN_OK;
}

extern "C" int bsn_fmr_set_mode(bsn_fmr* m, int inplace) {
    if (!m || (inplace != 0 && inplace != 1)) return set_err(BSN_ERR_ARG, "bad arguments");
    m->inplace = inplace;
    graph_drop(m->gexec, m->graph);
    return BSN_OK;
}

// Each level's plan tuned on the level's current state (the timing launches
// write the level's flux arrays and its Un buffer, which every level step
// rebuilds before reading).
extern "C" int bsn_fmr_autotune(bsn_fmr* m, char* report, int report_len) {
    if (!m) return set_err(BSN_ERR_ARG, "NULL fmr");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    std::string rep = "[";
    for (int l = 0; l < m->nlev; ++l) {
        if (l) rep += ", ";
        CHK(plan_autotune(*m->plans[l], rep));
    }
    rep += "]";
    graph_drop(m->gexec, m->graph);
    if (report && report_len > 0) snprintf(report, (size_t)report_len, "%s", rep.c_str());
    return BSN_OK;
}

extern "C" int bsn_fmr_prolong_all(bsn_fmr* m, int level) {
    if (!m || level < 1 || level >= m->nlev) return set_err(BSN_ERR_ARG, "bad level");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    CHK(fmr_prolong(m, level, 1, false));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr_prolong_fine_ghosts(bsn_fmr* m, int level) {
    if (!m || level < 1 || level >= m->nlev) return set_err(BSN_ERR_ARG, "bad level");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    CHK(fmr_prolong(m, level, 0, false));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr_restrict(bsn_fmr* m, int level) {
    if (!m || level < 1 || level >= m->nlev) return set_err(BSN_ERR_ARG, "bad level");
    HIPCHK(hipSetDevice(m->ctx->device));
    const FmrDev& f = m->hf[level - 1];
    const int64_t ncells = (int64_t)(f.hi[0] - f.lo[0] + 1) * (f.hi[1] - f.lo[1] + 1) * (f.hi[2] - f.lo[2] + 1);
    HIPCHK(launch_restrict(m->ctx->L(), m->d_all, m->d_f + (level - 1), ncells));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr3d_step(bsn_fmr* m, double dt_c) {
    if (!m) return set_err(BSN_ERR_ARG, "NULL fmr");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    HIPCHK(launch_set_scal(m->ctx->L(), m->lev[0]->h.scal, SCAL_DT, dt_c));
    CHK(fmr_advance(m, 0));
    return sync_ctx(m->ctx);
}

// ---------------------------------------------------------------------------
// FMR self-gravity (src/fmr_gravity.jl:40-133, src/fmr3d.jl:451-473)

static int64_t box_cells(const FmrDev& f) {
    return (int64_t)(f.hi[0] - f.lo[0] + 1) * (f.hi[1] - f.lo[1] + 1) * (f.hi[2] - f.lo[2] + 1);
}

static int64_t active_cells(const GridDev& h) { return (int64_t)h.nx * h.ny * h.nz; }

// Per-level phi buffers of the self-gravity solve.
static int fmr_sg_buffers(bsn_fmr* m) {
    if ((int)m->d_sgphi.size() != m->nlev) m->d_sgphi.assign(m->nlev, nullptr);
    for (int l = 0; l < m->nlev; ++l)
        if (!m->d_sgphi[l]) HIPCHK(hipMalloc((void**)&m->d_sgphi[l], active_cells(m->lev[l]->h) * 8));
    bsn_grid* root = m->lev[0];
    CHK(ensure_poisson(root, root->dx, root->dy, root->dz));
    return BSN_OK;
}

// solve_fmr_poisson over the chain of levels into phi[l]: the density of
// every level is restricted into its parent from the finest up (which
// changes the parent density inside the child's box, as
// restrict_density_to_coarse! does), the root's active density is solved
// with isolated boundaries, and phi is prolonged trilinearly down the chain;
// with G_grav != 1 every level's phi is then scaled.  With two levels this is
// the reference's function.  The device tables must be current.
static int fmr_sg_solve(bsn_fmr* m, double* const* phi, double G_grav) {
    const LaunchCtx L = m->ctx->L();
    for (int l = m->nlev - 1; l >= 1; --l)
        HIPCHK(launch_restrict(L, m->d_all, m->d_f + (l - 1), box_cells(m->hf[l - 1]), 1));
    bsn_grid* root = m->lev[0];
    HIPCHK(poisson_solve_state(*root->pois, root->h, 0, phi[0], L.stream));
    for (int l = 1; l < m->nlev; ++l) {
        const GridDev& pc = m->lev[l - 1]->h;
        const GridDev& cf = m->lev[l]->h;
        const FmrDev& f = m->hf[l - 1];
        const int nc[3] = {pc.nx, pc.ny, pc.nz}, nf[3] = {cf.nx, cf.ny, cf.nz};
        HIPCHK(launch_prolong_scalar(L, phi[l - 1], nc, phi[l], nf, f.lo, f.ratio));
    }
    if (G_grav != 1.0)
        for (int l = 0; l < m->nlev; ++l) HIPCHK(launch_scale(L, phi[l], active_cells(m->lev[l]->h), G_grav));
    return BSN_OK;
}

// The levels read the frozen phi of the solve (restored to their own phi,
// or none, by fmr_sg_end; the device tables follow at the next refresh).
static int fmr_sg_begin(bsn_fmr* m, std::vector<const double*>& saved) {
    CHK(fmr_sg_buffers(m));
    saved.clear();
    for (int l = 0; l < m->nlev; ++l) {
        saved.push_back(m->lev[l]->h.phi);
        m->lev[l]->h.phi = m->d_sgphi[l];
    }
    return fmr_refresh(m);
}

static void fmr_sg_end(bsn_fmr* m, const std::vector<const double*>& saved) {
    for (int l = 0; l < m->nlev && l < (int)saved.size(); ++l) m->lev[l]->h.phi = saved[l];
}

extern "C" int bsn_fmr3d_step_gravity(bsn_fmr* m, double dt_c, int self_gravity, double G_grav) {
    if (!m || (self_gravity != 0 && self_gravity != 1)) return set_err(BSN_ERR_ARG, "bad arguments");
    if (!self_gravity) return bsn_fmr3d_step(m, dt_c);
    if (!(G_grav == G_grav)) return set_err(BSN_ERR_ARG, "G_grav is NaN");
    HIPCHK(hipSetDevice(m->ctx->device));
    std::vector<const double*> saved;
    int r = fmr_sg_begin(m, saved);
    if (r == BSN_OK) r = fmr_sg_solve(m, m->d_sgphi.data(), G_grav);
    if (r == BSN_OK) {
        hipError_t e = launch_set_scal(m->ctx->L(), m->lev[0]->h.scal, SCAL_DT, dt_c);
        r = e == hipSuccess ? fmr_advance(m, 0) : set_err(BSN_ERR_HIP, "%s", hipGetErrorString(e));
    }
    if (r == BSN_OK) r = sync_ctx(m->ctx);
    fmr_sg_end(m, saved);
    return r;
}

extern "C" int bsn_solve_fmr_poisson(bsn_fmr* m, double* const* phi, int where) {
    if (!m || !phi) return set_err(BSN_ERR_ARG, "NULL argument");
    for (int l = 0; l < m->nlev; ++l)
        if (!phi[l]) return set_err(BSN_ERR_ARG, "NULL phi for level %d", l);
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    CHK(fmr_sg_buffers(m));
    std::vector<double*> dst(m->nlev);
    for (int l = 0; l < m->nlev; ++l) dst[l] = where == BSN_HOST ? m->d_sgphi[l] : phi[l];
    CHK(fmr_sg_solve(m, dst.data(), 1.0));
    if (where == BSN_HOST)
        for (int l = 0; l < m->nlev; ++l)
            HIPCHK(hipMemcpyAsync(phi[l], dst[l], active_cells(m->lev[l]->h) * 8, hipMemcpyDeviceToHost,
                                  m->ctx->stream));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr_restrict_density(bsn_fmr* m, int level) {
    if (!m || level < 1 || level >= m->nlev) return set_err(BSN_ERR_ARG, "bad level");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    HIPCHK(launch_restrict(m->ctx->L(), m->d_all, m->d_f + (level - 1), box_cells(m->hf[level - 1]), 1));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr_prolong_scalar(bsn_fmr* m, int level, const double* phi_c, int where_c,
                                      double* phi_f, int where_f) {
    if (!m || level < 1 || level >= m->nlev) return set_err(BSN_ERR_ARG, "bad level");
    if (!phi_c || !phi_f) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_sg_buffers(m));
    const GridDev& pc = m->lev[level - 1]->h;
    const GridDev& cf = m->lev[level]->h;
    const FmrDev& f = m->hf[level - 1];
    const int64_t ncn = active_cells(pc), nfn = active_cells(cf);
    // host arrays are staged through the self-gravity buffers of the two levels
    const double* src = phi_c;
    double* dst = where_f == BSN_HOST ? m->d_sgphi[level] : phi_f;
    if (where_c == BSN_HOST) {
        HIPCHK(hipMemcpyAsync(m->d_sgphi[level - 1], phi_c, ncn * 8, hipMemcpyHostToDevice, m->ctx->stream));
        src = m->d_sgphi[level - 1];
    }
    const int nc[3] = {pc.nx, pc.ny, pc.nz}, nf[3] = {cf.nx, cf.ny, cf.nz};
    HIPCHK(launch_prolong_scalar(m->ctx->L(), src, nc, dst, nf, f.lo, f.ratio));
    if (where_f == BSN_HOST)
        HIPCHK(hipMemcpyAsync(phi_f, dst, nfn * 8, hipMemcpyDeviceToHost, m->ctx->stream));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr_set_self_gravity(bsn_fmr* m, int on, double G_grav) {
    if (!m || (on != 0 && on != 1) || !(G_grav == G_grav)) return set_err(BSN_ERR_ARG, "bad arguments");
    m->sg_on = on;
    m->sg_G = G_grav;
    graph_drop(m->gexec, m->graph);
    return BSN_OK;
}

// Level CFL dts folded into the root: dt_0 = min(dt_0, r0 dt_1, r0 r1 dt_2, ...)
static int fmr_cfl(bsn_fmr* m, double cfl, double t_end) {
    const LaunchCtx L = m->ctx->L();
    double factor = 1.0;
    for (int l = 0; l < m->nlev; ++l) {
        Plan& P = *m->plans[l];
        CHK(plan_cfl(P, cfl, l == 0 ? t_end : 0.0));
        HIPCHK(launch_dt(L, P.d_grids, 1));
    }
    for (int l = 1; l < m->nlev; ++l) {
        factor *= (double)m->hf[l - 1].ratio;
        HIPCHK(launch_fold_dt(L, m->d_all, 0, l, factor));
    }
    return BSN_OK;
}

extern "C" int bsn_cfl_dt_fmr3d(bsn_fmr* m, double cfl, double* dt_out) {
    if (!m || !dt_out) return set_err(BSN_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(m->ctx->device));
    CHK(fmr_refresh(m));
    CHK(fmr_cfl(m, cfl, 0.0));
    HIPCHK(hipMemcpyAsync(dt_out, &m->lev[0]->h.scal[SCAL_DT], 8, hipMemcpyDeviceToHost, m->ctx->stream));
    return sync_ctx(m->ctx);
}

// One device-dt root step: the dt of the state at the start of the step
// (cfl_dt_fmr3d), then with self-gravity the composite solve, then the step.
static int fmr_device_step(bsn_fmr* m, double cfl, double t_end) {
    CHK(fmr_cfl(m, cfl, t_end));
    if (m->sg_on) CHK(fmr_sg_solve(m, m->d_sgphi.data(), m->sg_G));
    CHK(fmr_advance(m, 0));
    HIPCHK(launch_advance_t(m->ctx->L(), m->d_all, 1));
    return BSN_OK;
}

extern "C" int bsn_fmr_run(bsn_fmr* m, double cfl, double t_end, int nsteps, int use_graph) {
    if (!m || nsteps < 0) return set_err(BSN_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(m->ctx->device));
    if (m->sg_on) {
        // the Poisson solve's FFT plans are launched eagerly: no graph
        std::vector<const double*> saved;
        int r = fmr_sg_begin(m, saved);
        if (r == BSN_OK) {
            hipError_t e = launch_set_scal(m->ctx->L(), m->lev[0]->h.scal, SCAL_T, 0.0);
            if (e != hipSuccess) r = set_err(BSN_ERR_HIP, "%s", hipGetErrorString(e));
        }
        for (int s = 0; s < nsteps && r == BSN_OK; ++s) r = fmr_device_step(m, cfl, t_end);
        if (r == BSN_OK) r = sync_ctx(m->ctx);
        fmr_sg_end(m, saved);
        return r;
    }
    CHK(fmr_refresh(m));
    const LaunchCtx L = m->ctx->L();
    HIPCHK(launch_set_scal(L, m->lev[0]->h.scal, SCAL_T, 0.0));
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) CHK(fmr_device_step(m, cfl, t_end));
        return sync_ctx(m->ctx);
    }
    graph_drop(m->gexec, m->graph);
    HIPCHK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
    int r = fmr_device_step(m, cfl, t_end);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(L.stream, &gr);
    if (r != BSN_OK) { if (gr) (void)hipGraphDestroy(gr); return r; }
    if (e != hipSuccess) return set_err(BSN_ERR_HIP, "capture: %s", hipGetErrorString(e));
    m->graph = gr;
    HIPCHK(hipGraphInstantiate(&m->gexec, m->graph, nullptr, nullptr, 0));
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(m->gexec, L.stream));
    return sync_ctx(m->ctx);
}

extern "C" int bsn_fmr_time(bsn_fmr* m, double* t, double* dt) {
    if (!m) return set_err(BSN_ERR_ARG, "NULL fmr");
    HIPCHK(hipSetDevice(m->ctx->device));
    double s[SCAL_COUNT];
    HIPCHK(hipMemcpy(s, m->lev[0]->h.scal, sizeof s, hipMemcpyDeviceToHost));
    if (t) *t = s[SCAL_T];
    if (dt) *dt = s[SCAL_DT];
    return BSN_OK;
}
