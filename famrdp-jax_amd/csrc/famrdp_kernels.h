// Device-side data structures shared by the kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "famrdp_math.h"

namespace famrdp {

// One block of the multi-block state as the kernels see it.  Arrays are the
// reference layout: q (5, ni, nj, nk), kxyz (3, 3, ni, nj, nk) and vol
// (ni, nj, nk), C order with k unit stride; var/component stride is N.
struct BlockDesc {
  double* q;            // state (read at stage 1, updated in place at stage 3)
  const double* kxyz;
  const double* vol;
  double* q1;           // stage-1 result (workspace)
  double* q2;           // stage-2 result (workspace)
  double* G;            // viscous flux tensor (9, N) (workspace, direct path)
  int32_t ni, nj, nk;
  int32_t wrap_mod;     // 1: tile kernels wrap coordinates with a modulo (set by plan_create)
  int64_t N;            // ni*nj*nk (< 2^31 for blocks of a plan)
  double mu0;           // Sutherland mu0 of this block (ensembles of cases)
  double dt_scale;      // this block's time step = dt * dt_scale
  // adjoint workspace (null unless the plan was created for reverse sweeps)
  double* qh;           // post-halo/BC state recomputed for the reverse sweep (5, N)
  double* la;           // stage cotangents, ping (5, N)
  double* lb;           // stage cotangents, pong (5, N)
  double* l0;           // cotangent of the post-halo/BC state (5, N)
  double* fs;           // inviscid face reverse scratch (3 axes x 10, N)
  double* vs;           // viscous reverse scratch (16, N)
};

// One ghost cell written by the halo exchange: the final writer (dict order)
// of target cell `tc` of block `tb` reads source cell `sc` of block `sb` from
// the pre-exchange state.
struct HaloEntry {
  int32_t tb, tc, sb, sc;
};

enum BcKind : int32_t {
  BC_WALL_ADIABATIC = 0,
  BC_WALL_ISOTHERMAL = 1,
  BC_MIRROR_NORMAL = 2,    // symmetry plane / slip wall: flip the face-axis momentum
  BC_FARFIELD = 3,
  BC_EXTRAPOLATE = 4,      // extension: ghost layers copy the first interior layer
};

struct BcEntry {
  int32_t block, face, kind, pad_;
  double T_wall;
  double q_inf[5];
};

// Work tile of the batched (multi-block) launches: `len` consecutive cells of
// block `block` starting at flat cell index `start`.
struct Tile {
  int32_t block, start;
};

}  // namespace famrdp
