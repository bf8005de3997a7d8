// Test-only host build of the fused stage kernels' cell indexing
// (famrdp-jax_amd/csrc/famrdp_tiles.h: the slot maps of FusedGeom,
// fused_wrap_mode, tile_cell_rel, tile_cell), so the CPU tests can check, for
// every tile of a block and every thread, that the cells the kernels load are
// the periodic (roll) image of the cells the tile needs and lie in the block,
// and that every LDS slot the kernels address lies in its buffer and holds the
// cell it is read for.  Compiled with --cuda-host-only; never loaded by the
// product.
#include <cstdint>

#include "famrdp_tiles.h"

using namespace famrdp;

namespace {

// the periodic image of (i, j, k) in an ni x nj x nk block, computed with
// 64-bit modular arithmetic (independently of the kernels' wrap forms)
int64_t ref_index(int64_t i, int64_t j, int64_t k, int64_t ni, int64_t nj, int64_t nk) {
  i = ((i % ni) + ni) % ni;
  j = ((j % nj) + nj) % nj;
  k = ((k % nk) + nk) % nk;
  return (i * nj + j) * nk + k;
}

// the kernels' staging / G-cell index (fused_cells_fill): the cell at local
// coordinates l of the tile at (i0, j0, k0) under wrap mode w
int64_t rel_index(int w, Loc3 l, int i0, int j0, int k0, int ni, int nj, int nk) {
  if (w == WRAP_MOD) return tile_cell_rel<WRAP_MOD>(l.x + 2, l.y + 2, l.z + 2, i0, j0, k0, ni, nj, nk);
  if (w == WRAP_NONE) return tile_cell_rel<WRAP_NONE>(l.x + 2, l.y + 2, l.z + 2, i0, j0, k0, ni, nj, nk);
  return tile_cell_rel<WRAP_ONCE>(l.x + 2, l.y + 2, l.z + 2, i0, j0, k0, ni, nj, nk);
}

bool same(Loc3 a, Loc3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

int64_t own_index(int w, int i, int j, int k, int ni, int nj, int nk) {
  if (w == WRAP_MOD) return tile_cell<WRAP_MOD>(i, j, k, ni, nj, nk);
  if (w == WRAP_NONE) return tile_cell<WRAP_NONE>(i, j, k, ni, nj, nk);
  return tile_cell<WRAP_ONCE>(i, j, k, ni, nj, nk);
}

struct Counts {
  int64_t bad = 0, checked = 0, none_tiles = 0, tiles = 0;
};

template <int TI, int TJ, int TK>
void check_shape(int ni, int nj, int nk, Counts& c) {
  using T = FusedGeom<TI, TJ, TK>;
  using C = FusedCells<TI, TJ, TK>;
  const int64_t N = (int64_t)ni * nj * nk;
  const int wm = tile_wrap_mod(ni, nj, nk);
  auto check = [&](int64_t got, int64_t want) {
    ++c.checked;
    if (got != want || got < 0 || got >= N) ++c.bad;
  };
  // a slot read for cell l must lie in its buffer and belong to l
  auto pv_read = [&](Loc3 l) {
    const int s = T::pv_slot(l.x, l.y, l.z);
    ++c.checked;
    if (s < 0 || s >= T::NPV || !same(T::pv_cell(s), l)) ++c.bad;
  };
  auto g_read = [&](Loc3 l) {
    const int s = T::g_slot(l.x, l.y, l.z);
    ++c.checked;
    if (s < 0 || s >= T::NG || !same(T::g_cell(s), l)) ++c.bad;
  };
  // slot maps: pv_slot and g_slot invert pv_cell and g_cell
  for (int h = 0; h < T::NPV; ++h) {
    const Loc3 l = T::pv_cell(h);
    ++c.checked;
    if (T::pv_slot(l.x, l.y, l.z) != h) ++c.bad;
  }
  for (int h = 0; h < T::NG; ++h) {
    const Loc3 l = T::g_cell(h);
    ++c.checked;
    if (T::g_slot(l.x, l.y, l.z) != h) ++c.bad;
    // viscous-tensor phase: the G cell and its six neighbours' primitives
    pv_read(l);
    pv_read(Loc3{l.x + 1, l.y, l.z}); pv_read(Loc3{l.x - 1, l.y, l.z});
    pv_read(Loc3{l.x, l.y + 1, l.z}); pv_read(Loc3{l.x, l.y - 1, l.z});
    pv_read(Loc3{l.x, l.y, l.z + 1}); pv_read(Loc3{l.x, l.y, l.z - 1});
  }
  for (int t = 0; t < T::NT; ++t) {   // divergence: G of the own cell's six neighbours
    const int x = t / (TJ * TK), y = (t / TK) % TJ, z = t % TK;
    g_read(Loc3{x + 1, y, z}); g_read(Loc3{x - 1, y, z});
    g_read(Loc3{x, y + 1, z}); g_read(Loc3{x, y - 1, z});
    g_read(Loc3{x, y, z + 1}); g_read(Loc3{x, y, z - 1});
  }
  // every tile of the block (origins as the tile tables of a plan)
  for (int i0 = 0; i0 < ni; i0 += TI)
    for (int j0 = 0; j0 < nj; j0 += TJ)
      for (int k0 = 0; k0 < nk; k0 += TK) {
        const int w = fused_wrap_mode<TI, TJ, TK>(wm, i0, j0, k0, ni, nj, nk);
        ++c.tiles;
        if (w == WRAP_NONE) ++c.none_tiles;
        for (int t = 0; t < T::NT; ++t) {
          for (int it = 0; it < C::NIT; ++it) {
            const int h = t + it * T::NT;
            const Loc3 l = T::pv_cell(h < T::NPV ? h : T::NPV - 1);
            check(rel_index(w, l, i0, j0, k0, ni, nj, nk), ref_index(i0 + l.x, j0 + l.y, k0 + l.z, ni, nj, nk));
          }
          for (int it = 0; it < C::NGIT; ++it) {
            const int h = t + it * T::NT;
            const Loc3 l = T::g_cell(h < T::NG ? h : T::NG - 1);
            check(rel_index(w, l, i0, j0, k0, ni, nj, nk), ref_index(i0 + l.x, j0 + l.y, k0 + l.z, ni, nj, nk));
          }
          const int x = t / (TJ * TK), y = (t / TK) % TJ, z = t % TK;
          check(own_index(w, i0 + x, j0 + y, k0 + z, ni, nj, nk), ref_index(i0 + x, j0 + y, k0 + z, ni, nj, nk));
        }
      }
}

}  // namespace

extern "C" {

// Checks the tile shapes of the fused kernels on one ni x nj x nk block.
// out: [mismatches or out-of-block indices, values checked, tiles, tiles
// taking the no-wrap path]
void hc_tile_index_check(int ni, int nj, int nk, int64_t* out) {
  Counts c;
  check_shape<8, 4, 8>(ni, nj, nk, c);
  check_shape<4, 8, 16>(ni, nj, nk, c);
  check_shape<8, 4, 16>(ni, nj, nk, c);
  check_shape<4, 8, 8>(ni, nj, nk, c);
  check_shape<8, 8, 4>(ni, nj, nk, c);
  check_shape<4, 4, 16>(ni, nj, nk, c);
  check_shape<8, 8, 8>(ni, nj, nk, c);
  out[0] = c.bad;
  out[1] = c.checked;
  out[2] = c.tiles;
  out[3] = c.none_tiles;
}

}  // extern "C"
