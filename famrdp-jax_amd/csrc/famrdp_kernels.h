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
  double* l0;           // cotangent of the post-halo/BC state (5, N): the caller's cotangent array
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

// Doubles in global memory.  Pointers read from a device table (the batched
// launches' block descriptors) are generic to the compiler, which then loads
// through them with flat instructions (64-bit vector addresses, counted as
// LDS accesses as well); a cast to these types gives global loads and stores
// with a scalar base and a 32-bit vector offset.
typedef __attribute__((address_space(1))) const double gmem_cdouble;
typedef __attribute__((address_space(1))) double gmem_double;

// periodic neighbours along an axis of n cells
__device__ __forceinline__ int wrap_m(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrap_p(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// primitives of cell c of a (5, N) state
__device__ __forceinline__ void load_prim(const double* __restrict__ q, int64_t N, int64_t c, double g,
                                          double pv[5]) {
  double qc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) qc[v] = q[v * N + c];
  cons_to_prim(qc, g, pv);
}

__device__ __forceinline__ void load_prim_fast(const double* __restrict__ q, int64_t N, int64_t c, double g,
                                               double pv[5]) {
  double qc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) qc[v] = q[v * N + c];
  cons_to_prim_fast(qc, g, pv);
}

// XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so map
// consecutive ids of one XCD onto consecutive tiles (neighbouring tiles share
// halo lines in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int wg, int nwg) {
  constexpr int X = 8;
  const int q = nwg / X, r = nwg % X;
  const int xcd = wg % X, local = wg / X;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

}  // namespace famrdp
