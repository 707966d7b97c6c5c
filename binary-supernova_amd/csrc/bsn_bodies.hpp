// bsn_bodies.hpp -- gas/body couplings and gas diagnostics as device
// reductions over a grid's active cells (bsn_bodies.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"

namespace bsn {

enum { BQ_FORCE = 0, BQ_ACCRETE = 1, BQ_DIAG = 2 };

// Accumulators per block: force 3 per body, accretion 4 (one body per
// blockIdx.y), diagnostics 9.
#define BQ_NACC_FORCE (3 * BSN_MAX_BH)
#define BQ_NACC_ACCRETE 4
#define BQ_NACC_DIAG 9
#define BQ_WAVES 4           // waves (rows in flight) per block
#define BQ_MAX_BLOCKS 2048
// partial sums: the largest of 3 * BSN_MAX_BH force and 4 * BSN_MAX_BH accretion
// accumulators, each over BQ_MAX_BLOCKS blocks
#define BQ_PART_DOUBLES (4 * BSN_MAX_BH * BQ_MAX_BLOCKS)

// What a reduction needs besides the state: the bodies and the geometry.
// Active cell ii (1-based) has centre x0 + (ii - 0.5) dx.
struct BodyQ {
    int n;
    int pad;
    double x0, y0, z0, dx, dy, dz, dV;
    double gamma, rho_sink_min;
    double px[BSN_MAX_BH], py[BSN_MAX_BH], pz[BSN_MAX_BH];
    double mass[BSN_MAX_BH], eps[BSN_MAX_BH];
    double rs[BSN_MAX_BH], rate[BSN_MAX_BH];
    // accretion: the box of active cells (0-based first index, extent) that
    // can lie inside each body's sink radius
    int lo[BSN_MAX_BH][3], ext[BSN_MAX_BH][3];
};

// part needs BQ_PART_DOUBLES doubles; out receives the sums
// (force: 3n, accretion: 4n, diagnostics: 9), all on the stream.
hipError_t launch_body_force(hipStream_t s, int cus, const GridDev& G, const BodyQ& Q, double* part,
                             double* out);
hipError_t launch_body_accrete(hipStream_t s, int cus, const GridDev& G, const BodyQ& Q, double* part,
                               double* out);
hipError_t launch_body_diag(hipStream_t s, int cus, const GridDev& G, const BodyQ& Q, double* part,
                            double* out);

}  // namespace bsn
