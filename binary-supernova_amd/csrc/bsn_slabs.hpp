// bsn_slabs.hpp -- slab packing of a grid's state into file layouts (bsn_slabs.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"
#include "../../include/bsn.h"

namespace bsn {

// pack = 1: state -> buf; pack = 0: buf -> state.  Slab [k0, k0 + nk) of
// active k (BSN_LAYOUT_FIELDS, part = field) or of padded k (BSN_LAYOUT_REF).
hipError_t launch_slab(hipStream_t s, int cus, const GridDev& G, int pack, int layout, int part,
                       int k0, int nk, double* buf);

}  // namespace bsn
