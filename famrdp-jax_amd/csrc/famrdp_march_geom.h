// Geometry of the i-marching stage kernel (famrdp_march.h): LDS plane
// layouts, the halo cells each thread stages, the face list of a plane, the
// periodic wrap of tile columns.  Host-callable, so the index arithmetic can
// be checked on the CPU.
#pragma once

#include <hip/hip_runtime.h>

namespace famrdp {

// One j- or k-face of a plane: axis, left cell a along it (unwrapped block
// coordinate), the other in-plane coordinate oc, the P slot of the cell two
// before the face and the slot stride along the axis; ok = the face exists
// (left cell in [1, n-3], column inside the block).
struct MarchFace {
  bool isj, ok;
  int a, oc, base, str;
};

// What thread t of a tile at (j0, k0) handles for the whole march: its tile
// column (tj, tk) at block coordinates (gj, gk) (unwrapped; own = inside the
// block), the column's offset in a plane at its periodic image, its P and G
// slots, whether it takes the j / k inviscid terms (rhs/invscd.py:129-131);
// whether its column's low j-face and low k-face exist (fj, fk: the faces of
// rounds 0 and 1, see face_of); the halo cell hs of P it stages (-1: none)
// with its plane offset and P slot, and, for hs < NHG (hg), the G cross-halo
// cell of the same column.
struct MarchThread {
  int tj, tk, gj, gk, slot, gslot, hs, hslot, hgslot;
  bool own, jin, kin, fj, fk, hg;
  unsigned col, hcol;
};

// periodic image of coordinate x in [0, n) (x >= -2)
__host__ __device__ __forceinline__ int march_wrap(int x, int n) {
  const int r = x % n;
  return r < 0 ? r + n : r;
}

// A block of ni planes is marched in march_nseg(ni, L) segments of at most
// L planes, split evenly: segment `seg` owns planes [pA, pB).
__host__ __device__ __forceinline__ int march_nseg(int ni, int L) { return (ni + L - 1) / L; }
__host__ __device__ __forceinline__ void march_segment_bounds(int seg, int nseg, int ni, int& pA, int& pB) {
  pA = (int)((int64_t)seg * ni / nseg);
  pB = (int)((int64_t)(seg + 1) * ni / nseg);
}

// Tiles of one segment of an nj x nk plane for TJ x TK columns.  The plane is
// covered by TJ x TK tiles, except that the columns past the last full k
// tile (nk mod TK of them) go to transposed TK x TJ tiles when that takes
// fewer workgroups: a narrow remainder then costs a few strip tiles instead
// of a whole row of TJ x TK tiles that own a few columns each.  Tile t of a
// segment: the ntj x ntk main tiles first (k fastest), then the sj x sk
// strip tiles from column k1 on (k fastest).
struct MarchTiling {
  int ntj, ntk, k1, sj, sk;
  __host__ __device__ int count() const { return ntj * ntk + sj * sk; }
};
__host__ __device__ inline MarchTiling march_tiling(int TJ, int TK, int nj, int nk) {
  MarchTiling m;
  m.ntj = (nj + TJ - 1) / TJ;
  const int full = nk / TK, r = nk - full * TK;
  const int sj = (nj + TK - 1) / TK, sk = (r + TJ - 1) / TJ;
  if (TJ != TK && r > 0 && sj * sk < m.ntj) {
    m.ntk = full;
    m.k1 = full * TK;
    m.sj = sj;
    m.sk = sk;
  } else {
    m.ntk = (nk + TK - 1) / TK;
    m.k1 = nk;
    m.sj = 0;
    m.sk = 0;
  }
  return m;
}
// origin (j0, k0) of tile t of a segment, and whether it is a strip tile
// (transposed: TK x TJ columns)
__host__ __device__ inline void march_tile_of(const MarchTiling& m, int TJ, int TK, int t, int& j0, int& k0,
                                              int& tr) {
  const int nmain = m.ntj * m.ntk;
  if (t < nmain) {
    j0 = (t / m.ntk) * TJ;
    k0 = (t % m.ntk) * TK;
    tr = 0;
  } else {
    t -= nmain;
    j0 = (t / m.sk) * TK;
    k0 = m.k1 + (t % m.sk) * TJ;
    tr = 1;
  }
}

template <int TJ, int TK>
struct MarchGeom {
  static constexpr int NT = TJ * TK;                        // threads: one per tile column
  static constexpr int PJ = TJ + 4, PK = TK + 4, NP = PJ * PK;   // primitive plane (two halo rings)
  static constexpr int GJ = TJ + 2, GK = TK + 2, NGP = GJ * GK;  // viscous tensor plane (one ring)
  static constexpr int NFJ = (TJ + 1) * TK, NFK = TJ * (TK + 1), NFB = NFJ + NFK;   // faces of a plane
  static constexpr int NHG = 2 * (TJ + TK);                 // cross halo of G
  static constexpr int NHS = NHG + 4 + 2 * (TJ + TK);       // halo cells of P
  // doubles: P, G plane, face buffer, halo of the next plane, (u, v, w, T)
  // of the previous plane at the G cross halo
  static constexpr int NLDS = 6 * NP + 9 * NGP + 5 * NFB + 6 * NHS + 4 * NHG;
  // Thread roles.  Threads [0, NHG) also stage and compute G at the cross
  // halo cell of the same index; the other halo cells of P are staged by the
  // threads right after them (from T1 = NHG).  Faces go out in three rounds:
  // round 0 the low j-face of each thread's column, round 1 its low k-face,
  // round 2 the far-end layer (the j-faces past the last row, the k-faces
  // past the last column: NFL = TJ + TK faces) on the last wave (threads from
  // TF), which stages no halo cell.
  static constexpr int T1 = NHG;
  static constexpr int NFR = 3;
  static constexpr int NFL = TJ + TK;
  static constexpr int TF = NT - 64;
  static_assert(NT % 64 == 0 && NFL <= 64 && T1 + (NHS - NHG) <= TF && NFB == 2 * NT + NFL, "tile shape");

  __host__ __device__ static constexpr int ps(int lj, int lk) { return (lj + 2) * PK + lk + 2; }
  __host__ __device__ static constexpr int gs(int lj, int lk) { return (lj + 1) * GK + lk + 1; }
  // local (j, k) of halo cell h of P: the G cross halo (j = -1, j = TJ,
  // k = -1, k = TK), the four corners of the first ring, the second ring
  __host__ __device__ static constexpr void halo(int h, int& lj, int& lk) {
    if (h < TK) { lj = -1; lk = h; return; }
    h -= TK;
    if (h < TK) { lj = TJ; lk = h; return; }
    h -= TK;
    if (h < TJ) { lj = h; lk = -1; return; }
    h -= TJ;
    if (h < TJ) { lj = h; lk = TK; return; }
    h -= TJ;
    if (h < 4) { lj = h < 2 ? -1 : TJ; lk = (h & 1) ? TK : -1; return; }
    h -= 4;
    if (h < TK) { lj = -2; lk = h; return; }
    h -= TK;
    if (h < TK) { lj = TJ + 1; lk = h; return; }
    h -= TK;
    if (h < TJ) { lj = h; lk = -2; return; }
    h -= TJ;
    lj = h; lk = TK + 1;
  }
  // halo cell staged by thread t, or -1
  __host__ __device__ static constexpr int halo_of(int t) {
    return t < NHG ? t : ((t >= T1 && t < T1 + (NHS - NHG)) ? NHG + (t - T1) : -1);
  }
  // face of round r handled by thread t, or -1
  __host__ __device__ static constexpr int face_of(int t, int r) {
    if (r == 0) return t;                                     // j-face below row tj at column tk
    if (r == 1) return NFJ + (t / TK) * (TK + 1) + t % TK;    // k-face left of column tk in row tj
    const int u = t - TF;
    if (u < 0 || u >= NFL) return -1;
    return u < TK ? TJ * TK + u : NFJ + (u - TK) * (TK + 1) + TK;
  }
  __host__ __device__ static inline MarchThread thread(int t, int j0, int k0, int nj, int nk) {
    MarchThread m{};
    m.tj = t / TK;
    m.tk = t % TK;
    m.gj = j0 + m.tj;
    m.gk = k0 + m.tk;
    m.own = m.gj < nj && m.gk < nk;
    m.col = (unsigned)(march_wrap(m.gj, nj) * nk + march_wrap(m.gk, nk));
    m.slot = ps(m.tj, m.tk);
    m.gslot = gs(m.tj, m.tk);
    m.jin = m.own && m.gj >= 2 && m.gj <= nj - 3;
    m.kin = m.own && m.gk >= 2 && m.gk <= nk - 3;
    m.fj = m.gk < nk && m.gj >= 2 && m.gj <= nj - 2;
    m.fk = m.gj < nj && m.gk >= 2 && m.gk <= nk - 2;
    m.hs = halo_of(t);
    m.hg = t < NHG;
    m.hcol = 0;
    m.hslot = 0;
    m.hgslot = 0;
    if (m.hs >= 0) {
      int lj = 0, lk = 0;
      halo(m.hs, lj, lk);
      m.hcol = (unsigned)(march_wrap(j0 + lj, nj) * nk + march_wrap(k0 + lk, nk));
      m.hslot = ps(lj, lk);
      m.hgslot = m.hg ? gs(lj, lk) : 0;
    }
    return m;
  }
  // face f of the plane of a tile at (j0, k0) in an nj x nk plane: j-face
  // f < NFJ at (f / TK, f % TK) lies between local rows f / TK - 1 and
  // f / TK; k-face NFJ + r (TK + 1) + c between local columns c - 1 and c of
  // row r (written with selects: one instruction stream for both axes)
  __host__ __device__ static inline MarchFace face(int f, int j0, int k0, int nj, int nk) {
    MarchFace m{};
    m.isj = f < NFJ;
    const int fj = f / TK, lkj = f % TK;
    const int g2 = f - NFJ;
    const int ljk = g2 / (TK + 1), fk = g2 - ljk * (TK + 1);
    m.a = m.isj ? j0 + fj - 1 : k0 + fk - 1;
    m.oc = m.isj ? k0 + lkj : j0 + ljk;
    m.ok = m.isj ? (m.a >= 1 && m.a <= nj - 3 && m.oc < nk) : (m.a >= 1 && m.a <= nk - 3 && m.oc < nj);
    m.base = m.isj ? ps(fj - 2, lkj) : ps(ljk, fk - 2);
    m.str = m.isj ? PK : 1;
    return m;
  }
};

}  // namespace famrdp
