// Reverse of one RK3 stage (famrdp_reverse.hip, its own translation unit):
// the viscous and inviscid face reverse, then the per-cell gather, over the
// plan's adjoint block table and cell tiles.
#pragma once

#include <hip/hip_runtime.h>

#include "famrdp_kernels.h"

namespace famrdp {

void launch_stage_vjp(const BlockDesc* blocks, const Tile* tiles, int ntiles, int stage, const Gas& gas, double dt,
                      int lim, hipStream_t st);

}  // namespace famrdp
