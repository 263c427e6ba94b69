// mgp_geometry.h -- metric assembly on the GPU (mgp_geometry.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mgp2d.h"
#include "mgp_common.h"

namespace mgp {

// Throws MGP_ERR_ARG for an unknown kind or non-finite parameters.
void check_mapping(const mgp_mapping *m);
// J (cells), Jxx, Jxy (x-faces), Jyx, Jyy (y-faces) of the mapping, in the
// level layout of g.
void launch_metric(hipStream_t s, const mgp_mapping &m, const Grid &g, double *J, double *Jxx, double *Jxy,
                   double *Jyx, double *Jyy);

}  // namespace mgp
