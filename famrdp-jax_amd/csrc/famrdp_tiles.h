// Tile geometry of the fused and split stage kernels: the periodic wrap of
// tile coordinates into a block and the cell index (three wrap modes), and
// the LDS slot maps of a fused tile.  Host-callable, so the index arithmetic
// can be checked on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace famrdp {

// Periodic (roll) wrap of tile coordinates into a block and the cell index.
// Tiles start inside the block (origin <= n - 1) and the cells a tile kernel
// touches lie at most 2 cells before and T + 1 past its origin along an axis
// of tile extent T, so in a block whose padded sizes are all >= WRAP_ONCE_MIN
// > T a coordinate lies in [-n, 2n) and one conditional add or subtract wraps
// it; written as unsigned minima it takes two full-rate instructions and no
// branch.  With nj*nk < 2^24 the index then comes from 24-bit multiplies.
// Other blocks (BlockDesc::wrap_mod) take a modulo; the choice is uniform per
// block.  Tiles of the fused kernels whose cells all lie inside the block
// take no wrap at all (uniform per tile).  Plans hold blocks of fewer than
// 2^31 cells, so indices are 32-bit.
constexpr int WRAP_ONCE_MIN = 18;
constexpr int WRAP_ONCE = 0, WRAP_MOD = 1, WRAP_NONE = 2;   // tile_cell modes

struct Tile3 {
  int32_t block, i0, j0, k0;
};

// local coordinates of a cell relative to a tile's origin
struct Loc3 {
  int x, y, z;
};

template <int TI, int TJ, int TK>
struct FusedGeom {
  static_assert(TI < WRAP_ONCE_MIN && TJ < WRAP_ONCE_MIN && TK < WRAP_ONCE_MIN, "tile longer than the single-wrap bound");
  static constexpr int NT = TI * TJ * TK;
  static constexpr int B0J = TJ + 2, B0K = TK + 2;
  static constexpr int NR0 = (TI + 2) * B0J * B0K;      // box with a one-cell halo
  static constexpr int NR1 = 2 * TJ * TK;               // second i layer  [xe][y][z]
  static constexpr int NR2 = 2 * TI * TK;               // second j layer  [ye][x][z]
  static constexpr int NR3 = 2 * TI * TJ;               // second k layer  [ze][x][y]
  static constexpr int NPV = NR0 + NR1 + NR2 + NR3;
  static constexpr int NG = NT + 2 * (TJ * TK + TI * TK + TI * TJ);   // tile + cross halo 1
  static constexpr int F0 = (TI + 1) * TJ * TK, F1 = TI * (TJ + 1) * TK, F2 = TI * TJ * (TK + 1);
  static constexpr int NF = F0 > F1 ? (F0 > F2 ? F0 : F2) : (F1 > F2 ? F1 : F2);
  // primitives (5) and temperature per staged cell, then G or the faces
  static constexpr int NLDS = 6 * NPV + (9 * NG > 5 * NF ? 9 * NG : 5 * NF);
  // re-staging form: G overwrites the primitives (which are staged again for
  // the faces), so LDS holds the larger of the three phases only
  static constexpr int NLDS_RS_A = 6 * NPV > 9 * NG ? 6 * NPV : 9 * NG;
  static constexpr int NLDS_RS = NLDS_RS_A > 5 * NPV + 5 * NF ? NLDS_RS_A : 5 * NPV + 5 * NF;
  static constexpr int NPVS = NPV;   // primitive slots (stride of a variable)

  // slot of the cell (lx, ly, lz) whose coordinate along D lies in
  // [-2, T_D + 1] and whose other two coordinates lie in the tile
  template <int D>
  __device__ static __forceinline__ int axis_slot(int lx, int ly, int lz) {
    if constexpr (D == 0) {
      const bool out = (unsigned)(lx + 1) > (unsigned)(TI + 1);
      return out ? NR0 + ((lx < 0 ? 0 : 1) * TJ + ly) * TK + lz : ((lx + 1) * B0J + (ly + 1)) * B0K + (lz + 1);
    } else if constexpr (D == 1) {
      const bool out = (unsigned)(ly + 1) > (unsigned)(TJ + 1);
      return out ? NR0 + NR1 + ((ly < 0 ? 0 : 1) * TI + lx) * TK + lz : ((lx + 1) * B0J + (ly + 1)) * B0K + (lz + 1);
    } else {
      const bool out = (unsigned)(lz + 1) > (unsigned)(TK + 1);
      return out ? NR0 + NR1 + NR2 + ((lz < 0 ? 0 : 1) * TI + lx) * TJ + ly
                 : ((lx + 1) * B0J + (ly + 1)) * B0K + (lz + 1);
    }
  }

  // (written with selects only: no divergent branches in the face loops)
  __host__ __device__ static constexpr int pv_slot(int x, int y, int z) {
    const bool ox = (unsigned)(x + 1) > (unsigned)(TI + 1), oy = (unsigned)(y + 1) > (unsigned)(TJ + 1);
    const bool oz = (unsigned)(z + 1) > (unsigned)(TK + 1);
    const int s_box = ((x + 1) * B0J + (y + 1)) * B0K + (z + 1);
    const int s_x = NR0 + ((x < 0 ? 0 : 1) * TJ + y) * TK + z;
    const int s_y = NR0 + NR1 + ((y < 0 ? 0 : 1) * TI + x) * TK + z;
    const int s_z = NR0 + NR1 + NR2 + ((z < 0 ? 0 : 1) * TI + x) * TJ + y;
    return ox ? s_x : (oy ? s_y : (oz ? s_z : s_box));
  }
  // local coordinates of LDS primitive slot h (inverse of pv_slot)
  __host__ __device__ static constexpr Loc3 pv_cell(int h) {
    if (h < NR0) return Loc3{h / (B0J * B0K) - 1, (h / B0K) % B0J - 1, h % B0K - 1};
    h -= NR0;
    if (h < NR1) return Loc3{h / (TJ * TK) ? TI + 1 : -2, (h / TK) % TJ, h % TK};
    h -= NR1;
    if (h < NR2) return Loc3{(h / TK) % TI, h / (TI * TK) ? TJ + 1 : -2, h % TK};
    h -= NR2;
    return Loc3{(h / TJ) % TI, h % TJ, h / (TI * TJ) ? TK + 1 : -2};
  }
  __host__ __device__ static constexpr int g_slot(int x, int y, int z) {
    const bool ox = (unsigned)x >= (unsigned)TI, oy = (unsigned)y >= (unsigned)TJ, oz = (unsigned)z >= (unsigned)TK;
    const int s_in = (x * TJ + y) * TK + z;
    const int s_x = NT + ((x < 0 ? 0 : 1) * TJ + y) * TK + z;
    const int s_y = NT + 2 * TJ * TK + ((y < 0 ? 0 : 1) * TI + x) * TK + z;
    const int s_z = NT + 2 * TJ * TK + 2 * TI * TK + ((z < 0 ? 0 : 1) * TI + x) * TJ + y;
    return ox ? s_x : (oy ? s_y : (oz ? s_z : s_in));
  }
  // local coordinates of G slot h (inverse of g_slot)
  __host__ __device__ static constexpr Loc3 g_cell(int h) {
    if (h < NT) return Loc3{h / (TJ * TK), (h / TK) % TJ, h % TK};
    h -= NT;
    if (h < 2 * TJ * TK) return Loc3{h / (TJ * TK) ? TI : -1, (h / TK) % TJ, h % TK};
    h -= 2 * TJ * TK;
    if (h < 2 * TI * TK) return Loc3{(h / TK) % TI, h / (TI * TK) ? TJ : -1, h % TK};
    h -= 2 * TI * TK;
    return Loc3{(h / TJ) % TI, h % TJ, h / (TI * TJ) ? TK : -1};
  }
};

// Cell indices a thread of the fused kernel loads from: its staging slots
// t, t + NT, ... (indices past the last slot clamped to it, the values
// unused), the G cells beyond the tile it computes, and its own tile cell
// (wrapped into the block when outside it).  Computed once per tile; the
// wrap form (fused_wrap_mode) is chosen per tile by a uniform branch.
template <int TI, int TJ, int TK>
struct FusedCells {
  using T = FusedGeom<TI, TJ, TK>;
  static constexpr int NIT = (T::NPV + T::NT - 1) / T::NT;
  static constexpr int NGIT = (T::NG + T::NT - 1) / T::NT;
  int stage[NIT];
  int gh[NGIT > 1 ? NGIT - 1 : 1];
  int c;
};

__host__ __device__ inline int tile_wrap_mod(int ni, int nj, int nk) {
  return (ni < WRAP_ONCE_MIN || nj < WRAP_ONCE_MIN || nk < WRAP_ONCE_MIN || (int64_t)nj * nk >= (1 << 24)) ? 1 : 0;
}

__host__ __device__ __forceinline__ int wrap_once(int i, int n) {   // i in [-n, 2n)
  const unsigned a = (unsigned)i < (unsigned)(i + n) ? (unsigned)i : (unsigned)(i + n);
  const unsigned b = a - (unsigned)n;
  return (int)(a < b ? a : b);
}

__host__ __device__ __forceinline__ int wrap_mod(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

// 24-bit unsigned multiply (both operands < 2^24); the host form serves the
// host-compiled index checks
__host__ __device__ __forceinline__ unsigned umul24(unsigned a, unsigned b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __umul24(a, b);
#else
  return (a & 0xffffffu) * (b & 0xffffffu);
#endif
}

// index of cell (i, j, k) of the block, each coordinate wrapped (W is one of
// WRAP_ONCE, WRAP_MOD, WRAP_NONE; bool true / false select MOD / ONCE)
template <int W>
__host__ __device__ __forceinline__ int tile_cell(int i, int j, int k, int ni, int nj, int nk) {
  if constexpr (W == WRAP_MOD) {
    return (wrap_mod(i, ni) * nj + wrap_mod(j, nj)) * nk + wrap_mod(k, nk);
  } else if constexpr (W == WRAP_ONCE) {
    return (int)(umul24((unsigned)wrap_once(i, ni), (unsigned)(nj * nk)) +
                 umul24((unsigned)wrap_once(j, nj), (unsigned)nk)) +
           wrap_once(k, nk);
  } else {
    return (int)(umul24((unsigned)i, (unsigned)(nj * nk)) + umul24((unsigned)j, (unsigned)nk)) + k;
  }
}

// cell index of the cell at (px - 2, py - 2, pz - 2) from the tile origin
// (i0, j0, k0); px, py, pz >= 0
template <int W>
__host__ __device__ __forceinline__ int tile_cell_rel(int px, int py, int pz, int i0, int j0, int k0, int ni, int nj,
                                                      int nk) {
  if constexpr (W == WRAP_NONE) {
    // (i0 - 2 + px) * nj*nk + ... : the origin part is uniform across the tile
    const int base = tile_cell<WRAP_NONE>(i0 - 2, j0 - 2, k0 - 2, ni, nj, nk);
    return base + (int)(umul24((unsigned)px, (unsigned)(nj * nk)) + umul24((unsigned)py, (unsigned)nk)) + pz;
  } else {
    return tile_cell<W>(i0 - 2 + px, j0 - 2 + py, k0 - 2 + pz, ni, nj, nk);
  }
}

// Wrap mode of a fused tile: the block's (WRAP_MOD or WRAP_ONCE), or none
// when every cell the tile touches (2 before its origin to T + 1 past it, per
// axis) lies in the block.  Uniform per tile.
template <int TI, int TJ, int TK>
__host__ __device__ __forceinline__ int fused_wrap_mode(int wrap_mod_block, int i0, int j0, int k0, int ni, int nj,
                                                        int nk) {
  if (wrap_mod_block) return WRAP_MOD;
  const bool inside = i0 >= 2 && j0 >= 2 && k0 >= 2 && i0 + TI + 2 <= ni && j0 + TJ + 2 <= nj && k0 + TK + 2 <= nk;
  return inside ? WRAP_NONE : WRAP_ONCE;
}

}  // namespace famrdp
