// Test-only host build of the i-marching stage kernel's geometry
// (famrdp-jax_amd/csrc/famrdp_march_geom.h: MarchGeom::thread, ::face,
// ::halo_of, ::face_of and the LDS slot maps), so the CPU tests can check,
// for every workgroup of a block, that every LDS slot the kernel writes lies
// in its buffer and is written once per step, that every slot it reads was
// written in the same step with the cell (or viscous tensor, or face) the
// read needs -- the periodic (roll) image for the viscous terms, the
// unwrapped cell for the faces -- that the faces' cells lie in the block,
// and that the updates cover every cell of the block exactly once.  The
// workgroups are the ones the kernel launches (march_tiling: TJ x TK tiles
// and, for a narrow k remainder, transposed TK x TJ strip tiles).
// Compiled with --cuda-host-only; never loaded by the product.
#include <cstdint>
#include <vector>

#include "famrdp_march_geom.h"

using namespace famrdp;

namespace {

struct Acc {
  int64_t bad = 0, checked = 0;
  void ok(bool c) {
    ++checked;
    if (!c) ++bad;
  }
};

int wrap(int x, int n) { return ((x % n) + n) % n; }

// thread roles of a tile shape: every halo cell staged once, every face once
template <int TJ, int TK>
void check_roles(Acc& acc) {
  using M = MarchGeom<TJ, TK>;
  acc.ok(M::NLDS * 8 <= 160 * 1024);
  {
    std::vector<int> hcount(M::NHS, 0), fcount(M::NFB, 0);
    for (int t = 0; t < M::NT; ++t) {
      const int h = M::halo_of(t);
      if (h >= 0) {
        acc.ok(h < M::NHS);
        if (h < M::NHS) ++hcount[h];
      }
      for (int r = 0; r < M::NFR; ++r) {
        const int f = M::face_of(t, r);
        if (f >= 0) {
          acc.ok(f < M::NFB);
          if (f < M::NFB) ++fcount[f];
        }
      }
    }
    for (int h = 0; h < M::NHS; ++h) acc.ok(hcount[h] == 1);
    for (int f = 0; f < M::NFB; ++f) acc.ok(fcount[f] == 1);
  }
}

// one workgroup: a TJ x TK tile at (j0, k0) over every segment
template <int TJ, int TK>
void check_tile(int ni, int nj, int nk, int L, int j0, int k0, Acc& acc, std::vector<int>& cover) {
  using M = MarchGeom<TJ, TK>;
    {
      // P: block column (j * nk + k) held by each slot; G plane likewise
      std::vector<int64_t> P(M::NP, -1), G(M::NGP, -1);
      // face buffer: (axis, left cell, other coordinate) of each slot
      std::vector<int64_t> F(M::NFB, -1);
      auto colof = [&](int j, int k) { return (int64_t)wrap(j, nj) * nk + wrap(k, nk); };
      auto putP = [&](int sl, int64_t c) {
        acc.ok(sl >= 0 && sl < M::NP);
        if (sl < 0 || sl >= M::NP) return;
        acc.ok(P[sl] == -1);
        P[sl] = c;
      };
      auto getP = [&](int sl, int64_t want) {
        acc.ok(sl >= 0 && sl < M::NP && P[sl] == want);
      };
      // a. P writes
      for (int t = 0; t < M::NT; ++t) {
        const MarchThread th = M::thread(t, j0, k0, nj, nk);
        acc.ok(th.col < (unsigned)(nj * nk) && (int64_t)th.col == colof(th.gj, th.gk));
        putP(th.slot, th.col);
        if (th.hs >= 0) {
          int lj = 0, lk = 0;
          M::halo(th.hs, lj, lk);
          acc.ok(th.hcol < (unsigned)(nj * nk) && (int64_t)th.hcol == colof(j0 + lj, k0 + lk));
          putP(th.hslot, th.hcol);
        }
      }
      // b. G at the column, e. G at the cross halo: P neighbours; G writes
      for (int t = 0; t < M::NT; ++t) {
        const MarchThread th = M::thread(t, j0, k0, nj, nk);
        getP(th.slot, colof(th.gj, th.gk));
        getP(th.slot + M::PK, colof(th.gj + 1, th.gk));
        getP(th.slot - M::PK, colof(th.gj - 1, th.gk));
        getP(th.slot + 1, colof(th.gj, th.gk + 1));
        getP(th.slot - 1, colof(th.gj, th.gk - 1));
        acc.ok(th.gslot >= 0 && th.gslot < M::NGP && G[th.gslot] == -1);
        if (th.gslot >= 0 && th.gslot < M::NGP) G[th.gslot] = th.col;
        if (th.hg) {
          int lj = 0, lk = 0;
          M::halo(th.hs, lj, lk);
          const int hj = j0 + lj, hk = k0 + lk;
          getP(th.hslot, colof(hj, hk));
          getP(th.hslot + M::PK, colof(hj + 1, hk));
          getP(th.hslot - M::PK, colof(hj - 1, hk));
          getP(th.hslot + 1, colof(hj, hk + 1));
          getP(th.hslot - 1, colof(hj, hk - 1));
          acc.ok(th.hgslot >= 0 && th.hgslot < M::NGP && G[th.hgslot] == -1);
          if (th.hgslot >= 0 && th.hgslot < M::NGP) G[th.hgslot] = th.hcol;
        }
      }
      // e. faces: four unwrapped cells in the block from P; kxyz offsets
      for (int t = 0; t < M::NT; ++t)
        for (int r = 0; r < M::NFR; ++r) {
          const int f = M::face_of(t, r);
          if (f < 0) continue;
          const MarchFace fc = M::face(f, j0, k0, nj, nk);
          // rounds 0 and 1: the column's own low j- and k-face, which the
          // kernel addresses from the column's slot and plane offset
          if (r < 2) {
            const MarchThread th = M::thread(t, j0, k0, nj, nk);
            acc.ok(fc.isj == (r == 0) && fc.ok == (r == 0 ? th.fj : th.fk));
            acc.ok(f == (r == 0 ? t : M::NFJ + th.tj * (TK + 1) + th.tk));
            if (fc.ok) {
              acc.ok(fc.base == th.slot - (r == 0 ? 2 * M::PK : 2) && fc.str == (r == 0 ? M::PK : 1));
              const int64_t cl = r == 0 ? (int64_t)fc.a * nk + fc.oc : (int64_t)fc.oc * nk + fc.a;
              acc.ok(cl == (int64_t)th.col - (r == 0 ? nk : 1));
            }
          } else {
            acc.ok(t >= M::TF && M::halo_of(t) < 0);
          }
          if (!fc.ok) continue;
          for (int s4 = 0; s4 < 4; ++s4) {
            const int c = fc.a - 1 + s4;          // unwrapped along the face axis
            const int j = fc.isj ? c : fc.oc, k = fc.isj ? fc.oc : c;
            acc.ok(j >= 0 && j < nj && k >= 0 && k < nk);
            getP(fc.base + s4 * fc.str, (int64_t)j * nk + k);
          }
          F[f] = (fc.isj ? 0 : 1) + 2 * ((int64_t)fc.a * 65536 + fc.oc);
        }
      // c. divergence reads of G, d. inviscid reads of the face buffer
      for (int t = 0; t < M::NT; ++t) {
        const MarchThread th = M::thread(t, j0, k0, nj, nk);
        if (!th.own) continue;
        acc.ok(G[th.gslot + M::GK] == colof(th.gj + 1, th.gk));
        acc.ok(G[th.gslot - M::GK] == colof(th.gj - 1, th.gk));
        acc.ok(G[th.gslot + 1] == colof(th.gj, th.gk + 1));
        acc.ok(G[th.gslot - 1] == colof(th.gj, th.gk - 1));
        if (th.jin) {
          const int lo = th.tj * TK + th.tk;
          acc.ok(F[lo] == 0 + 2 * ((int64_t)(th.gj - 1) * 65536 + th.gk));
          acc.ok(F[lo + TK] == 0 + 2 * ((int64_t)th.gj * 65536 + th.gk));
        }
        if (th.kin) {
          const int lo = M::NFJ + th.tj * (TK + 1) + th.tk;
          acc.ok(F[lo] == 1 + 2 * ((int64_t)(th.gk - 1) * 65536 + th.gj));
          acc.ok(F[lo + 1] == 1 + 2 * ((int64_t)th.gk * 65536 + th.gj));
        }
        // every plane of every segment: updated by this column once
        const int nseg = march_nseg(ni, L);
        for (int sg = 0; sg < nseg; ++sg) {
          int pA = 0, pB = 0;
          march_segment_bounds(sg, nseg, ni, pA, pB);
          acc.ok(pB - pA >= 1 && pB - pA <= L);
          for (int s = pA - 1; s <= pB; ++s)
            if (s > pA) ++cover[((int64_t)(s - 1) * nj + th.gj) * nk + th.gk];
        }
      }
    }
}

// every workgroup of a launch with main shape TJ x TK, as march_tiling and
// march_tile_of lay them out; returns the number of strip tiles
template <int TJ, int TK>
int check(int ni, int nj, int nk, int L, Acc& acc, std::vector<int>& cover) {
  check_roles<TJ, TK>(acc);
  check_roles<TK, TJ>(acc);
  acc.ok(MarchGeom<TJ, TK>::NLDS == MarchGeom<TK, TJ>::NLDS);
  const MarchTiling mt = march_tiling(TJ, TK, nj, nk);
  int strips = 0;
  for (int t = 0; t < mt.count(); ++t) {
    int j0 = 0, k0 = 0, tr = 0;
    march_tile_of(mt, TJ, TK, t, j0, k0, tr);
    acc.ok(j0 >= 0 && j0 < nj && k0 >= 0 && k0 < nk);
    if (tr) {
      ++strips;
      check_tile<TK, TJ>(ni, nj, nk, L, j0, k0, acc, cover);
    } else {
      check_tile<TJ, TK>(ni, nj, nk, L, j0, k0, acc, cover);
    }
  }
  return strips;
}

}  // namespace

extern "C" void hc_march_geom_check(int ni, int nj, int nk, int64_t* out) {
  Acc acc;
  int64_t cells = (int64_t)ni * nj * nk, wrong_cover = 0, strips = 0;
  const int shapes[5][3] = {{16, 16, 32}, {8, 32, 20}, {8, 32, 32}, {8, 32, 52}, {8, 32, 10}};
  for (const auto& sh : shapes) {
    std::vector<int> cover((size_t)cells, 0);
    if (sh[0] == 16) strips += check<16, 16>(ni, nj, nk, sh[2], acc, cover);
    else strips += check<8, 32>(ni, nj, nk, sh[2], acc, cover);
    for (int64_t c = 0; c < cells; ++c) wrong_cover += cover[c] != 1;
  }
  out[0] = acc.bad;
  out[1] = acc.checked;
  out[2] = wrong_cover;
  out[3] = strips;
}
