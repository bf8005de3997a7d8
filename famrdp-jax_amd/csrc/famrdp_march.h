// i-marching stage kernel: one launch per RK3 stage (solver/step.py:45-53),
// the same operator as the fused tile kernel (rhs/invscd.py:53-131,
// rhs/viscous.py:30-108), computed with the same expressions in the same
// order, so both give the same bits.
//
// A workgroup owns a column of TJ x TK cells (one thread per (j, k) column)
// over a segment [pA, pB) of planes along i, the slowest axis, and walks it
// one plane per step.  Step s, for s = pA-1 .. pB:
//
//   a. The primitives and temperature of plane s go to the LDS plane P: the
//      tile's columns and two halo rings (the second one without corners),
//      at periodic (roll) coordinates.  They were converted in the previous
//      step from loads issued a step before that.
//   b. The i-face between planes s-1 and s of the thread's own column (its
//      left state was reconstructed in the previous step; the right state
//      and the next face's left state share one limited slope per
//      variable), and the viscous flux tensor G of plane s at the column
//      (j/k differences from P, i differences from the column's registers).
//   c. The viscous divergence of plane s-1: j/k differences of G(s-1) from
//      the LDS G plane, the i difference from G(s) and G(s-2) in registers.
//   d. The inviscid terms of plane s-1 (i from the two i-faces around it,
//      j and k from the LDS face buffer filled in the previous step) and the
//      Shu-Osher combination; plane s-1 is written.
//   e. G(s) replaces G(s-1) in LDS, with G of the tile's one-cell cross halo;
//      the MUSCL-2 + Roe faces of plane s along j and k go to the face
//      buffer (each thread its column's low j-face, then its low k-face,
//      both with offsets fixed by the column; the last wave the far-end
//      layer); the halo of plane s+1 is converted for the next step.
//
// Loads are issued a phase or more before their use where registers allow:
// the column's metrics of plane s before the barrier of phase a (and kept for
// the next step's divergence), the update operands of plane s-1 and the
// cross halo's metrics in phase b, the state of plane s+2 a step ahead.
//
// Each cell's primitives, viscous tensor and faces are computed once per
// segment, apart from the tile's halo rings in j and k and the two extra
// planes at each segment start.  Cell coordinates are fixed per thread for
// the whole march (one plane stride apart from step to step).  Tiles that
// reach past the block in j or k (partial tiles, small blocks) stage and
// compute G at the periodic images of those columns, which is what the
// roll of the viscous terms reads; only columns inside the block write.  A
// narrow k remainder goes to transposed TK x TJ strip tiles (march_tiling),
// so that it does not take a row of TJ x TK tiles owning a few columns each.
// Loads and stores go through buffer resources with 32-bit byte offsets, so
// a block's kxyz must stay under 4 GiB (march_fits).
#pragma once

#include "famrdp_march_geom.h"

namespace famrdp {

// primitives (rho, u, v, w, p) and temperature of one cell, as the fused
// kernel stages them
DF void march_prim(const double q[5], double g, double R, double pv[6]) {
  cons_to_prim_fast(q, g, pv);
  pv[5] = temperature_fast(pv, R);
}

// The arrays a march reads and writes, as buffer resources: a load or store
// takes the element's 32-bit byte offset in a vector register and the
// variable's offset v N (bytes) in a scalar register, so no address is formed
// per access (global loads from base + v N formed a 64-bit vector address for
// every variable but the first: 1 % of the stage's time).  Every array of a
// block that march_fits accepts spans fewer than 2^32 bytes, which is the
// resource's record count; the last word is the gfx9-family setting for
// untyped (raw) access (32-bit data format, stride 0).
typedef unsigned march_u2 __attribute__((__vector_size__(8)));

DF __amdgpu_buffer_rsrc_t march_rsrc(const double* p, int64_t n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0, (int)(unsigned)(n * 8), 0x00020000);
}

// element i (< 2^29) of variable v: byte offset i * 8 + soff, soff = v N * 8
DF double ldx(__amdgpu_buffer_rsrc_t r, unsigned i, unsigned soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i << 3, (int)soff, 0));
}

DF void stx(__amdgpu_buffer_rsrc_t r, unsigned i, unsigned soff, double x) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(march_u2, x), r, i << 3, (int)soff, 0);
}

DF void march_load5(__amdgpu_buffer_rsrc_t r, unsigned nb, unsigned off, double x[5]) {
#pragma unroll
  for (int v = 0; v < 5; ++v) x[v] = ldx(r, off, v * nb);
}

// kxyz of the largest block a marching launch may address (9 N doubles
// under 4 GiB)
__host__ __device__ __forceinline__ bool march_fits(int64_t N) { return N * 72 < (int64_t(1) << 32); }

// Viscous flux tensor of the cell at P slot `sl` (rhs/viscous.py:30-108):
// j/k neighbours from P, i neighbours (u, v, w, T) from um (plane s-1) and up
// (plane s+1); kch = kxyz / 2 at the cell.  The fused kernel's expressions,
// with the 1/2 of each central difference moved onto the metrics: halving
// is exact, so every product d k of the gradients is the same number and
// the tensor has the same bits, for 12 multiplies fewer.
template <int TJ, int TK>
DF void march_visc(const double* sP, int sl, const double um[4], const double up[4], const double kch[9],
                   const Gas& gb, double g9[9]) {
  using M = MarchGeom<TJ, TK>;
  constexpr int NP = M::NP;
  const double pv[5] = {0.0, sP[NP + sl], sP[2 * NP + sl], sP[3 * NP + sl], 0.0};
  const double Tc = sP[5 * NP + sl];
  double d[3][4];
#pragma unroll
  for (int s = 0; s < 4; ++s) d[0][s] = up[s] - um[s];
#pragma unroll
  for (int l = 1; l < 3; ++l) {
    const int sp = l == 1 ? sl + M::PK : sl + 1;
    const int sm = l == 1 ? sl - M::PK : sl - 1;
    d[l][0] = sP[NP + sp] - sP[NP + sm];
    d[l][1] = sP[2 * NP + sp] - sP[2 * NP + sm];
    d[l][2] = sP[3 * NP + sp] - sP[3 * NP + sm];
    d[l][3] = sP[5 * NP + sp] - sP[5 * NP + sm];
  }
  viscous_flux_fast(pv, Tc, d, kch, gb, g9);
}

// One j- or k-face of the plane (rhs/invscd.py:53-97) into face buffer slot
// f: the four cells' primitives from P slots base + s str (s = 0..3), kxyz
// of the face's axis at the cells either side (plane offsets cl, cr; the
// axis's first row of kxyz at byte offset s0).
template <int TJ, int TK, int LIM>
DF void march_face(const double* sP, double* sF, int base, int str, __amdgpu_buffer_rsrc_t rk, unsigned cl,
                   unsigned cr, unsigned s0, unsigned nb, int f, double g) {
  using M = MarchGeom<TJ, TK>;
  double p4[4][5];
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) {
#pragma unroll
    for (int v = 0; v < 5; ++v) p4[s4][v] = sP[v * M::NP + base + s4 * str];
  }
  double kl[3], kr[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    kl[m] = ldx(rk, cl, s0 + m * nb);
    kr[m] = ldx(rk, cr, s0 + m * nb);
  }
  double Fa[5];
  face_flux_prim<LIM>(p4[0], p4[1], p4[2], p4[3], kl, kr, g, Fa);
#pragma unroll
  for (int v = 0; v < 5; ++v) sF[v * M::NFB + f] = Fa[v];
}

// A thread's own low j- or k-face (rounds 0 and 1): the right-hand cell is
// the thread's column, whose kxyz / 2 of the plane (krh, rows s0 / nb ..) the
// thread holds in registers; the left-hand row (plane offset cl) is loaded
// and halved, which is exact, so the face has the bits of march_face.
template <int TJ, int TK, int LIM>
DF void march_face_own(const double* sP, double* sF, int base, int str, __amdgpu_buffer_rsrc_t rk, unsigned cl,
                       unsigned s0, unsigned nb, const double krh[3], int f, double g) {
  using M = MarchGeom<TJ, TK>;
  double p4[4][5];
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) {
#pragma unroll
    for (int v = 0; v < 5; ++v) p4[s4][v] = sP[v * M::NP + base + s4 * str];
  }
  double klh[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) klh[m] = 0.5 * ldx(rk, cl, s0 + m * nb);
  double Fa[5];
  face_flux_prim_h<LIM>(p4[0], p4[1], p4[2], p4[3], klh, krh, g, Fa);
#pragma unroll
  for (int v = 0; v < 5; ++v) sF[v * M::NFB + f] = Fa[v];
}

// The far-end layer of the plane at offset `op` (face_of round 2: j-faces
// past the tile's last row, k-faces past its last column), for the threads
// of the last wave; j-face f (< NFJ) at (f / TK, f % TK) lies between local
// rows f / TK - 1 and f / TK, k-face NFJ + r (TK + 1) + c between local
// columns c - 1 and c of row r.  Faces exist between cells 1 .. n-2 of their
// axis (left cell a in [1, n-3]) and never read a wrapped cell.
template <int TJ, int TK, int LIM>
DF void march_far_faces(const BlockDesc& b, __amdgpu_buffer_rsrc_t rk, unsigned nb, int j0, int k0, unsigned op,
                        const double* sP, double* sF, double g, int t) {
  using M = MarchGeom<TJ, TK>;
  const int f = M::face_of(t, 2);
  if (f < 0) return;
  const int nj = b.nj, nk = b.nk;
  const MarchFace fc = M::face(f, j0, k0, nj, nk);
  if (!fc.ok) return;
  const bool isj = fc.isj;
  const unsigned cl = op + (isj ? (unsigned)fc.a * nk + fc.oc : (unsigned)fc.oc * nk + fc.a) + (isj ? 3u : 6u) * (unsigned)b.N;
  const unsigned cr = cl + (isj ? (unsigned)nk : 1u);
  march_face<TJ, TK, LIM>(sP, sF, fc.base, fc.str, rk, cl, cr, 0, nb, f, g);
}

// v, hidden from the optimiser: what is computed from it inside the march
// loop is computed there, not hoisted out and held in registers
__device__ __forceinline__ unsigned march_opaque(unsigned v) {
  asm volatile("" : "+v"(v));
  return v;
}

// x as computed (rounded), hidden from the optimiser: the fused kernel
// stages primitives, viscous tensors and face fluxes through LDS before
// using them, so the compiler never contracts the multiply that produced a
// value into the add that consumes it; values used here straight from
// registers are held the same way, which keeps the bits equal
DF void march_keep(double& x) { asm volatile("" : "+v"(x)); }
template <int K>
DF void march_keep(double (&x)[K]) {
#pragma unroll
  for (int i = 0; i < K; ++i) march_keep(x[i]);
}

template <int TJ, int TK, int LIM>
__device__ __forceinline__ void march_segment(double* lds, const BlockDesc& b, int j0, int k0, int pA, int pB,
                                              const double* __restrict__ qin, const double* q0, double* qout,
                                              const Gas& gas, double dt, int stage) {
  using M = MarchGeom<TJ, TK>;
  double* sP = lds;                          // [6][NP]  rho u v w p T of plane s
  double* sG = sP + 6 * M::NP;               // [9][NGP] G of plane s-1 (then s)
  double* sF = sG + 9 * M::NGP;              // [5][NFB] j/k faces of plane s-1 (then s)
  double* sH = sF + 5 * M::NFB;              // [6][NHS] halo of the next plane
  double* sU = sH + 6 * M::NHS;              // [4][NHG] u v w T of plane s-1 at the G cross halo
  const int t = threadIdx.x;
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const unsigned si = (unsigned)nj * (unsigned)nk;
  const double g = gas.gamma;
  Gas gb = gas;
  gb.mu0 = b.mu0;
  dt = dt * b.dt_scale;
  const unsigned nb = (unsigned)N * 8u;     // bytes between variables
  const __amdgpu_buffer_rsrc_t rq = march_rsrc(qin, 5 * N), r0 = march_rsrc(q0, 5 * N),
                               ro = march_rsrc(qout, 5 * N), rk = march_rsrc(b.kxyz, 9 * N),
                               rv = march_rsrc(b.vol, N);

  // ---- fixed per thread for the whole march --------------------------------
  const MarchThread th = M::thread(t, j0, k0, nj, nk);
  const int tj = th.tj, tk = th.tk, slot = th.slot, gslot = th.gslot, hs = th.hs;
  const int hslot = th.hslot, hgslot = th.hgslot;
  const bool own = th.own, jin = th.jin, kin = th.kin, fj = th.fj, fk = th.fk, hg = th.hg;
  const unsigned col = th.col, hcol = th.hcol;
  auto plane = [&](int s) -> unsigned { return (unsigned)(s < 0 ? s + ni : (s >= ni ? s - ni : s)) * si; };

  // ---- warm-up: planes pA-2, pA-1 converted, pA loading ---------------------
  double pm[6], pc[6];
  {
    double qq[5], x[6];
    march_load5(rq, nb, plane(pA - 2) + col, qq);
    march_prim(qq, g, gb.R, pm);
    march_load5(rq, nb, plane(pA - 1) + col, qq);
    march_prim(qq, g, gb.R, pc);
    march_keep(pm);
    march_keep(pc);
    if (hg) {
      march_load5(rq, nb, plane(pA - 2) + hcol, qq);
      march_prim(qq, g, gb.R, x);
      sU[hs] = x[1];
      sU[M::NHG + hs] = x[2];
      sU[2 * M::NHG + hs] = x[3];
      sU[3 * M::NHG + hs] = x[5];
    }
    if (hs >= 0) {
      march_load5(rq, nb, plane(pA - 1) + hcol, qq);
      march_prim(qq, g, gb.R, x);
#pragma unroll
      for (int v = 0; v < 6; ++v) sH[v * M::NHS + hs] = x[v];
    }
  }
  double qn[5], qnh[5] = {0, 0, 0, 0, 0};
  march_load5(rq, nb, plane(pA) + col, qn);
  if (hs >= 0) march_load5(rq, nb, plane(pA) + hcol, qnh);

  double kpvh[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // kxyz / 2 of the column at plane s-1
  double pLn[5] = {0, 0, 0, 0, 0};       // left state of the column's next i-face
  double Fm[5] = {0, 0, 0, 0, 0};        // i-face below plane s-1
  double Gm2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // G(s-2) of the column

  for (int s = pA - 1; s <= pB; ++s) {
    const unsigned op = plane(s), om = plane(s - 1);
    const unsigned c_s = op + march_opaque(col), c_m = om + march_opaque(col);
    const bool faces = s >= pA && s < pB;            // plane s is owned: its j/k faces
    const bool upd = s > pA;                         // plane s-1 is owned: its update (s-1 < pB always)
    // the column's metrics of plane s, loaded once (issued before the barrier,
    // which covers their latency): the i-face takes its right-hand row from
    // them, and they are kept for the next step's i-face (left-hand row) and
    // divergence of G(s)
    double kk[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) kk[m] = ldx(rk, c_s, m * nb);
    // ---- a. P <- plane s ------------------------------------------------------
#pragma unroll
    for (int v = 0; v < 6; ++v) sP[v * M::NP + slot] = pc[v];
    if (hs >= 0) {
#pragma unroll
      for (int v = 0; v < 6; ++v) sP[v * M::NP + hslot] = sH[v * M::NHS + hs];
    }
    __syncthreads();

    // ---- b. i-face between planes s-1 and s; G of plane s at the column --------
    double pn[6];
    march_prim(qn, g, gb.R, pn);
    march_keep(pn);
    if (s + 2 <= pB + 1) march_load5(rq, nb, plane(s + 2) + march_opaque(col), qn);
    // the metrics halved: the i-face normal is their sum, and the central
    // differences of the viscous terms are taken without their 1/2
    double kkh[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) kkh[m] = 0.5 * kk[m];
    double Fi[5] = {0, 0, 0, 0, 0};
    {
      double pLs[5], pRs[5];
      muscl_cell<LIM>(pm, pc, pn, pLs, pRs);
      face_state_fence(pLn);
      face_state_fence(pRs);
      if (s >= pA && s - 1 >= 1 && s - 1 <= ni - 3) face_flux_lr_h(pLn, pRs, kpvh, kkh, g, Fi);
      march_keep(Fi);
#pragma unroll
      for (int v = 0; v < 5; ++v) pLn[v] = pLs[v];
    }
    // the update operands of plane s-1, issued before G(s) is computed
    double qs[5] = {0, 0, 0, 0, 0}, qq0[5] = {0, 0, 0, 0, 0}, vl = 1.0;
    if (upd && own) {
      vl = ldx(rv, c_m, 0);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        qs[v] = ldx(rq, c_m, v * nb);
        qq0[v] = stage == 1 ? 0.0 : ldx(r0, c_m, v * nb);
      }
    }
    double g9[9];
    {
      const double um[4] = {pm[1], pm[2], pm[3], pm[5]}, up[4] = {pn[1], pn[2], pn[3], pn[5]};
      march_visc<TJ, TK>(sP, slot, um, up, kkh, gb, g9);
      march_keep(g9);
    }
    // metrics of the G cross-halo cell of plane s, for phase e
    double kh[9];
    if (hg) {
      const unsigned ch = op + march_opaque(hcol);
#pragma unroll
      for (int m = 0; m < 9; ++m) kh[m] = ldx(rk, ch, m * nb);
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) {
      pm[v] = pc[v];
      pc[v] = pn[v];
    }

    // ---- c. divergence of G(s-1) ------------------------------------------------
    double vis[4] = {0.0, 0.0, 0.0, 0.0};
    if (upd && own) {
      const double* km = kpvh;   // the differences' 1/2 on the metrics, as in march_visc
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double res = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
          const int c = gidx(r, jj);
          const double* Gc = sG + c * M::NGP;
          const double d0 = g9[c] - Gm2[c];
          const double d1 = Gc[gslot + M::GK] - Gc[gslot - M::GK];
          const double d2 = Gc[gslot + 1] - Gc[gslot - 1];
          res = res + (d0 * km[0 * 3 + jj] + d1 * km[1 * 3 + jj] + d2 * km[2 * 3 + jj]);
        }
        vis[r] = res;
      }
    }
#pragma unroll
    for (int m = 0; m < 9; ++m) Gm2[m] = sG[m * M::NGP + gslot];    // G(s-1): G(s'-2) of the next step
#pragma unroll
    for (int m = 0; m < 9; ++m) kpvh[m] = kkh[m];

    // ---- d. update of plane s-1 -------------------------------------------------
    if (upd && own) {
      march_keep(vl);
      const double rvol = rcp(vl + 1e-300);
      double inv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (s - 1 >= 2 && s - 1 <= ni - 3) {
#pragma unroll
        for (int v = 0; v < 5; ++v) inv[v] = inv[v] + (-(Fi[v] - Fm[v]) * rvol);
      }
      if (jin) {
        const int lo = tj * TK + tk;
#pragma unroll
        for (int v = 0; v < 5; ++v) inv[v] = inv[v] + (-(sF[v * M::NFB + lo + TK] - sF[v * M::NFB + lo]) * rvol);
      }
      if (kin) {
        const int lo = M::NFJ + tj * (TK + 1) + tk;
#pragma unroll
        for (int v = 0; v < 5; ++v) inv[v] = inv[v] + (-(sF[v * M::NFB + lo + 1] - sF[v * M::NFB + lo]) * rvol);
      }
      double rhs[5];
      rhs[0] = inv[0] + 0.0;
      rhs[1] = inv[1] + vis[0];
      rhs[2] = inv[2] + vis[1];
      rhs[3] = inv[3] + vis[2];
      rhs[4] = inv[4] + vis[3];
      // loaded in phase b; held opaque here so that the combination below
      // compiles with the same contractions as the fused kernel's, which
      // loads these values right where it combines them
      march_keep(qs);
      march_keep(qq0);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        stx(ro, c_m, v * nb, rk_combine(stage, qq0[v], qs[v], dt, rhs[v]));
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) Fm[v] = Fi[v];
    __syncthreads();   // every read of G(s-1) and of the faces of plane s-1 done

    // ---- e. G(s) over G(s-1); G of the cross halo; faces of plane s; next halo --
#pragma unroll
    for (int m = 0; m < 9; ++m) sG[m * M::NGP + gslot] = g9[m];
    if (hs >= 0) {
      double x[6];
      march_prim(qnh, g, gb.R, x);     // plane s+1 at the halo cell
      march_keep(x);
      if (s + 2 <= pB + 1) march_load5(rq, nb, plane(s + 2) + march_opaque(hcol), qnh);
#pragma unroll
      for (int v = 0; v < 6; ++v) sH[v * M::NHS + hs] = x[v];
      if (hg) {
        const double um[4] = {sU[hs], sU[M::NHG + hs], sU[2 * M::NHG + hs], sU[3 * M::NHG + hs]};
        const double up[4] = {x[1], x[2], x[3], x[5]};
        double khh[9], g9h[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) khh[m] = 0.5 * kh[m];
        march_visc<TJ, TK>(sP, hslot, um, up, khh, gb, g9h);
#pragma unroll
        for (int m = 0; m < 9; ++m) sG[m * M::NGP + hgslot] = g9h[m];
        sU[hs] = sP[M::NP + hslot];
        sU[M::NHG + hs] = sP[2 * M::NP + hslot];
        sU[2 * M::NHG + hs] = sP[3 * M::NP + hslot];
        sU[3 * M::NHG + hs] = sP[5 * M::NP + hslot];
      }
    }
    if (faces) {
      // round 0: the column's low j-face (cells gj-2 .. gj+1 of column gk;
      // in the block wherever the face exists, so its plane offsets are
      // those of the column and the column nk before it)
      if (fj)
        march_face_own<TJ, TK, LIM>(sP, sF, slot - 2 * M::PK, M::PK, rk, c_s - (unsigned)nk, 3 * nb, nb, kpvh + 3, t, g);
      __builtin_amdgcn_sched_barrier(0);   // one face at a time: the rounds are not interleaved
      // round 1: the column's low k-face
      if (fk)
        march_face_own<TJ, TK, LIM>(sP, sF, slot - 2, 1, rk, c_s - 1u, 6 * nb, nb, kpvh + 6, M::NFJ + tj * (TK + 1) + tk, g);
      __builtin_amdgcn_sched_barrier(0);
      march_far_faces<TJ, TK, LIM>(b, rk, nb, j0, k0, op, sP, sF, g, (int)march_opaque((unsigned)t));
    }
    __syncthreads();   // G(s) and the faces of plane s in place; P free
  }
}

// One tile's march: TJ x TK columns, or TK x TJ for a strip tile (tr); both
// shapes use the same LDS layout size.
template <int TJ, int TK, int LIM>
__device__ __forceinline__ void march_tile(double* lds, const BlockDesc& b, int tr, int j0, int k0, int pA,
                                           int pB, const double* __restrict__ qin, double* qout, int stage,
                                           const Gas& gas, double dt) {
  static_assert(MarchGeom<TJ, TK>::NLDS == MarchGeom<TK, TJ>::NLDS, "transposed tile LDS");
  if constexpr (TJ == TK) {
    march_segment<TJ, TK, LIM>(lds, b, j0, k0, pA, pB, qin, b.q, qout, gas, dt, stage);
  } else {
    if (tr) march_segment<TK, TJ, LIM>(lds, b, j0, k0, pA, pB, qin, b.q, qout, gas, dt, stage);
    else march_segment<TJ, TK, LIM>(lds, b, j0, k0, pA, pB, qin, b.q, qout, gas, dt, stage);
  }
}

// Plain launch: grid = segments x march_tiling(...).count(), XCD-remapped.
// Segments split the block's planes evenly (lengths differ by at most one).
template <int TJ, int TK, int L, int LIM, int WPE>
__global__ void __launch_bounds__(TJ * TK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_stage_march(BlockDesc b, int stage, Gas gas, double dt) {
  __shared__ double lds[MarchGeom<TJ, TK>::NLDS];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const MarchTiling mt = march_tiling(TJ, TK, b.nj, b.nk);
  const int nt = mt.count(), seg = w / nt;
  int j0, k0, tr;
  march_tile_of(mt, TJ, TK, w - seg * nt, j0, k0, tr);
  int pA, pB;
  march_segment_bounds(seg, march_nseg(b.ni, L), b.ni, pA, pB);
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
  march_tile<TJ, TK, LIM>(lds, b, tr, j0, k0, pA, pB, qin, qout, stage, gas, dt);
}

// Batched launch over a tile table whose i0 is 2 x segment index + strip bit.
template <int TJ, int TK, int L, int LIM, int WPE>
__global__ void __launch_bounds__(TJ * TK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_stage_march_batched(const BlockDesc* __restrict__ blocks, const Tile3* __restrict__ tiles, int stage, Gas gas,
                      double dt) {
  __shared__ double lds[MarchGeom<TJ, TK>::NLDS];
  const Tile3 tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  const BlockDesc& b = blocks[tl.block];
  int pA, pB;
  march_segment_bounds(tl.i0 >> 1, march_nseg(b.ni, L), b.ni, pA, pB);
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
  march_tile<TJ, TK, LIM>(lds, b, tl.i0 & 1, tl.j0, tl.k0, pA, pB, qin, qout, stage, gas, dt);
}

}  // namespace famrdp
