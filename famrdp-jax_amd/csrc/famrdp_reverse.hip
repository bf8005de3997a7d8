// Reverse-mode stage kernels of the TVD-RK3 step.
//
// One RK stage  q_out = a q0 + b (q_s + dt L(q_s))  is reversed by two
// gather-form kernels (no atomics, deterministic):
//   k_vjp_faces_visc
//                 every cell x: cotangent of the viscous flux tensor G(x)
//                 from the divergence at its six neighbours, reverse sweep of
//                 G into d/d(u, v, w, T) and d/d(central differences); then
//                 every inviscid face whose left cell is x: cotangent of
//                 F*|n| from the two cells that difference it, Roe + Harten
//                 reverse sweep, stored as (d/d pL, d/d pR) at x;
//   k_vjp_cell    every cell y: gathers the face states, the MUSCL slopes of
//                 y-1, y, y+1 (limiter reverse), the central differences of
//                 its neighbours and T, maps the primitive cotangent to
//                 conservative variables and applies the RK coefficients.
// Functions cite the forward code they reverse in famrdp_jax (rhs/invscd.py,
// rhs/viscous.py, solver/step.py).  This file is its own translation unit,
// compiled with the default machine scheduler (the rest of the library uses
// the max-ilp one): these two kernels measured 2.5 % and 6.5 % faster with it,
// the marching stage kernel 2.4 % slower.
#include "famrdp_kernels.h"
#include "famrdp_adjoint.h"
#include "famrdp_reverse.h"

namespace famrdp {

// Cells are addressed with 32-bit element indices (an adjoint plan's blocks
// hold fewer than 2^29 cells: famrdp_plan_create refuses larger ones, whose 80
// doubles per cell of stage and adjoint buffers would exceed 300 GB), loads and stores through a 32-bit byte offset from
// a base that is the same for the whole workgroup -- the base stays in scalar
// registers and a cell costs one 32-bit offset, however many arrays it reads
// (global loads and stores: the pointers come from the device block table).
__device__ __forceinline__ double rd(const double* __restrict__ a, unsigned i) {
  return *(gmem_cdouble*)(reinterpret_cast<const char*>(a) + (i << 3));
}
__device__ __forceinline__ gmem_double& at(double* a, unsigned i) {
  return *(gmem_double*)(reinterpret_cast<char*>(a) + (i << 3));
}

__device__ __forceinline__ void rd_prim(const double* __restrict__ q, int64_t N, unsigned c, double g, double pv[5]) {
  double qc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) qc[v] = rd(q + v * N, c);
  cons_to_prim_fast(qc, g, pv);
}

// ---------------------------------------------------------------------------
// inviscid faces (rhs/invscd.py:53-131 reversed)
// ---------------------------------------------------------------------------
template <int LIM>
__device__ __forceinline__ void cell_vjp_faces(const BlockDesc& b, const double* __restrict__ qs,
                                               const double* __restrict__ lin, double c_mu, unsigned a, double g,
                                               const double pva[5], const double kka[9]) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const unsigned njk = (unsigned)nj * (unsigned)nk;
  const int k = (int)(a % (unsigned)nk), j = (int)((a / (unsigned)nk) % (unsigned)nj), i = (int)(a / njk);
  // the cell's own 1/vol and scaled cotangent, shared by its three faces
  double lma[5];
  const double rva = rcp_hd(rd(b.vol, a) + 1e-300);
#pragma unroll
  for (int v = 0; v < 5; ++v) lma[v] = c_mu * rd(lin + v * N, a);
  // axes k, j, i: the viscous part that runs first in the same thread ends
  // with the k-neighbours, so the k faces find them in cache; each axis writes
  // only its own face slots, so the order leaves every result unchanged
#pragma unroll
  for (int d = 2; d >= 0; --d) {
    const int ad = d == 0 ? i : (d == 1 ? j : k);
    const int nd = d == 0 ? ni : (d == 1 ? nj : nk);
    const unsigned sd = d == 0 ? njk : (d == 1 ? (unsigned)nk : 1u);
    double lpl[5] = {0, 0, 0, 0, 0}, lpr[5] = {0, 0, 0, 0, 0};
    const bool va = ad >= 2 && ad <= nd - 3, vb = ad + 1 >= 2 && ad + 1 <= nd - 3;
    if (ad >= 1 && ad <= nd - 3 && (va || vb)) {
      const unsigned a1 = a + sd;
      double lFa[5];
      const double ra = va ? rva : 0.0, rb = vb ? rcp_hd(rd(b.vol, a1) + 1e-300) : 0.0;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double ma = va ? lma[v] : 0.0, mb = vb ? c_mu * rd(lin + v * N, a1) : 0.0;
        lFa[v] = -ra * ma + rb * mb;
      }
      double pv[4][5];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (s == 1) {
#pragma unroll
          for (int v = 0; v < 5; ++v) pv[1][v] = pva[v];
          continue;
        }
        double qq[5];
        const unsigned cc = a + (unsigned)(s - 1) * sd;
#pragma unroll
        for (int v = 0; v < 5; ++v) qq[v] = rd(qs + v * N, cc);
        cons_to_prim_fast(qq, g, pv[s]);
      }
      double pL[5], pR[5];
      face_states<LIM>(pv[0], pv[1], pv[2], pv[3], pL, pR);
      double k0[3], k1[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        k0[m] = kka[d * 3 + m];
        k1[m] = rd(b.kxyz + (d * 3 + m) * N, a1);
      }
      double nu[3], nm;
      face_normal(k0, k1, nu, nm);
      double lF[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) lF[v] = lFa[v] * nm;
      roe_flux_prim_vjp(pL, pR, nu, g, lF, lpl, lpr);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      at(b.fs + (d * 10 + v) * N, a) = lpl[v];
      at(b.fs + (d * 10 + 5 + v) * N, a) = lpr[v];
    }
  }
}

// ---------------------------------------------------------------------------
// viscous flux tensor (rhs/viscous.py:86-108 and 30-84 reversed)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cell_vjp_visc(const BlockDesc& b, const double* __restrict__ qs,
                                              const double* __restrict__ lin, double c_mu, unsigned x,
                                              const Gas& gas, const double pv[5], const double kk[9]) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const unsigned sj = (unsigned)nk, si = (unsigned)nj * (unsigned)nk;
  const int k = (int)(x % sj), j = (int)((x / sj) % (unsigned)nj), i = (int)(x / si);
  const unsigned cm[3] = {x + (unsigned)(wrap_m(i, ni) - i) * si, x + (unsigned)(wrap_m(j, nj) - j) * sj,
                          x + (unsigned)(wrap_m(k, nk) - k)};
  const unsigned cp[3] = {x + (unsigned)(wrap_p(i, ni) - i) * si, x + (unsigned)(wrap_p(j, nj) - j) * sj,
                          x + (unsigned)(wrap_p(k, nk) - k)};
  // cotangent of G(x): the divergence at x - e_l uses +G(x)/2, at x + e_l uses -G(x)/2
  double lG[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const unsigned ym = cm[l], yp = cp[l];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double mm = c_mu * rd(lin + (r + 1) * N, ym), mp = c_mu * rd(lin + (r + 1) * N, yp);
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        lG[gidx(r, jj)] += 0.5 * (rd(b.kxyz + (l * 3 + jj) * N, ym) * mm - rd(b.kxyz + (l * 3 + jj) * N, yp) * mp);
    }
  }
  double d[3][4];
  const double T = temperature_fast(pv, gas.R);
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    double pp[5], pm[5];
    rd_prim(qs, N, cp[l], gas.gamma, pp);
    rd_prim(qs, N, cm[l], gas.gamma, pm);
    d[l][0] = 0.5 * (pp[1] - pm[1]);
    d[l][1] = 0.5 * (pp[2] - pm[2]);
    d[l][2] = 0.5 * (pp[3] - pm[3]);
    d[l][3] = 0.5 * (temperature_fast(pp, gas.R) - temperature_fast(pm, gas.R));
  }
  double lu = 0, lv = 0, lw = 0, lT = 0, ld[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  viscous_flux_vjp(pv, T, d, kk, gas.gamma, gas.R, b.mu0, gas.T0, gas.Ts, gas.Pr_lam, lG, lu, lv, lw, lT, ld);
  at(b.vs, x) = lu;
  at(b.vs + N, x) = lv;
  at(b.vs + 2 * N, x) = lw;
  at(b.vs + 3 * N, x) = lT;
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int s = 0; s < 4; ++s) at(b.vs + (4 + 4 * l + s) * N, x) = ld[l][s];
}

// ---------------------------------------------------------------------------
// per-cell gather, primitive -> conservative, RK coefficients
//   mode 0: l0 = c0 lin,  lout = c_next lin + g     (stage 3; lin is l0, each
//           cell's value read before it is written)
//   mode 1: l0 += c0 lin, lout = c_next lin + g     (stage 2)
//   mode 2: l0 += lin + g                           (stage 1)
// ---------------------------------------------------------------------------
template <int LIM>
__device__ __forceinline__ void cell_vjp_gather(const BlockDesc& b, const double* __restrict__ qs,
                                                const double* lin, double* lout, unsigned y, const Gas& gas,
                                                double c_next, double c0, int mode) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const unsigned sj = (unsigned)nk, si = (unsigned)nj * (unsigned)nk;
  const int k = (int)(y % sj), j = (int)((y / sj) % (unsigned)nj), i = (int)(y / si);
  // the cell's own state and primitives, converted once for the three axes
  double qy[5], pvy[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) qy[v] = rd(qs + v * N, y);
  cons_to_prim_fast(qy, gas.gamma, pvy);
  double lpv[5] = {0, 0, 0, 0, 0};
  // ---- inviscid: face states and MUSCL slopes along each axis
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double* fsd = b.fs + (int64_t)(d * 10) * N;
    const int yd = d == 0 ? i : (d == 1 ? j : k);
    const int nd = d == 0 ? ni : (d == 1 ? nj : nk);
    const unsigned sd = d == 0 ? si : (d == 1 ? sj : 1u);
    double fl[5], fr[5];   // cotangents of the left state of face (y, y+1), right state of (y-1, y)
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      fl[v] = rd(fsd + v * N, y);
      fr[v] = yd >= 1 ? rd(fsd + (5 + v) * N, y - sd) : 0.0;
      lpv[v] += fl[v];
      if (yd >= 1) lpv[v] += fr[v];
    }
    // slopes s(x) = lim(pv(x)-pv(x-1), pv(x+1)-pv(x)) for x = y-1, y, y+1
    double pv[5][5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      const int xd = yd + s - 2;
      if (s == 2)
#pragma unroll
        for (int v = 0; v < 5; ++v) pv[2][v] = pvy[v];
      else if (xd >= 0 && xd < nd) rd_prim(qs, N, y + (unsigned)(s - 2) * sd, gas.gamma, pv[s]);
      else
#pragma unroll
        for (int v = 0; v < 5; ++v) pv[s][v] = 0.0;
    }
#pragma unroll
    for (int o = -1; o <= 1; ++o) {
      const int xd = yd + o;
      if (xd < 1 || xd > nd - 2) continue;
      const unsigned xc = y + (unsigned)o * sd;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double ls = o == 0 ? 0.5 * fl[v] - 0.5 * fr[v]
                                 : 0.5 * rd(fsd + v * N, xc) - 0.5 * rd(fsd + (5 + v) * N, xc - sd);
        const double A = pv[o + 2][v] - pv[o + 1][v], B = pv[o + 3][v] - pv[o + 2][v];
        double lA, lB;
        limiter_vjp_t<LIM>(A, B, ls, lA, lB);
        if (o == 0) lpv[v] += lA - lB;
        else if (o == 1) lpv[v] -= lA;
        else lpv[v] += lB;
      }
    }
  }
  // ---- viscous: own uses plus the central differences of the neighbours
  const unsigned cm[3] = {y + (unsigned)(wrap_m(i, ni) - i) * si, y + (unsigned)(wrap_m(j, nj) - j) * sj,
                          y + (unsigned)(wrap_m(k, nk) - k)};
  const unsigned cp[3] = {y + (unsigned)(wrap_p(i, ni) - i) * si, y + (unsigned)(wrap_p(j, nj) - j) * sj,
                          y + (unsigned)(wrap_p(k, nk) - k)};
  double ls4[4] = {rd(b.vs, y), rd(b.vs + N, y), rd(b.vs + 2 * N, y), rd(b.vs + 3 * N, y)};
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int s = 0; s < 4; ++s)
      ls4[s] += 0.5 * rd(b.vs + (4 + 4 * l + s) * N, cm[l]) - 0.5 * rd(b.vs + (4 + 4 * l + s) * N, cp[l]);
  lpv[1] += ls4[0]; lpv[2] += ls4[1]; lpv[3] += ls4[2];
  temperature_vjp(pvy, gas.R, ls4[3], lpv[0], lpv[4]);
  double gq[5] = {0, 0, 0, 0, 0};
  cons_to_prim_vjp(qy, gas.gamma, lpv, gq);
  // ---- RK coefficients (solver/step.py:45-53 reversed)
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double li = rd(lin + v * N, y);
    gmem_double& l0 = at(b.l0 + v * N, y);
    if (mode == 2) {
      l0 += li + gq[v];
    } else {
      l0 = (mode == 0 ? 0.0 : l0) + c0 * li;
      at(lout + v * N, y) = c_next * li + gq[v];
    }
  }
}

// Stage selector for the reverse sweep: which buffers a stage reads/writes.
//   stage 3: qs = q2, lin = l0 (the cotangent of q_new, which the gather
//            replaces cell by cell with its q0 share: l0 = lin / 3), lout = lb
//   stage 2: qs = q1, lin = lb, lout = la
//   stage 1: qs = qh, lin = la
__device__ __forceinline__ void vjp_bufs(const BlockDesc& b, int stage, const double*& qs, const double*& lin,
                                         double*& lout) {
  qs = stage == 3 ? b.q2 : (stage == 2 ? b.q1 : b.qh);
  lin = stage == 3 ? b.l0 : (stage == 2 ? b.lb : b.la);
  lout = stage == 3 ? b.lb : (stage == 2 ? b.la : nullptr);
}

__device__ __forceinline__ double stage_cmu(int stage, double dt) {
  return stage == 3 ? (2.0 / 3.0) * dt : (stage == 2 ? 0.25 * dt : dt);
}

// The viscous part first: the face part reuses the neighbours' state and
// metrics it loaded (from registers or cache).  The scheduling barrier keeps
// the two parts' instructions apart (two waves per SIMD, as the face part
// alone needs: 207 registers with the default scheduler, no spills).
template <int LIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_vjp_faces_visc(const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles, int stage, Gas gas, double dt) {
  const Tile tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];   // neighbouring tiles on one XCD (shared L2)
  const BlockDesc& b = blocks[tl.block];
  const int64_t c = (int64_t)tl.start + threadIdx.x;
  if (c >= b.N) return;
  const double *qs, *lin;
  double* lout;
  vjp_bufs(b, stage, qs, lin, lout);
  const double cmu = stage_cmu(stage, dt * b.dt_scale);
  // the cell's primitives, converted once for both parts (and rounded once:
  // the same values reach every use), and its metrics, loaded once for both
  double pv[5], kk[9];
  rd_prim(qs, b.N, (unsigned)c, gas.gamma, pv);
#pragma unroll
  for (int m = 0; m < 9; ++m) kk[m] = rd(b.kxyz + m * b.N, (unsigned)c);
  cell_vjp_visc(b, qs, lin, cmu, (unsigned)c, gas, pv, kk);
  __builtin_amdgcn_sched_barrier(0);
  cell_vjp_faces<LIM>(b, qs, lin, cmu, (unsigned)c, gas.gamma, pv, kk);
}

// at most 128 VGPRs: this memory-bound gather wants four waves per SIMD
template <int LIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_vjp_cell(const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
                                                  int stage, Gas gas) {
  const Tile tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];   // neighbouring tiles on one XCD (shared L2)
  const BlockDesc& b = blocks[tl.block];
  const int64_t c = (int64_t)tl.start + threadIdx.x;
  if (c >= b.N) return;
  const double *qs, *lin;
  double* lout;
  vjp_bufs(b, stage, qs, lin, lout);
  const double c_next = stage == 3 ? 2.0 / 3.0 : 0.25;
  const double c0 = stage == 3 ? 1.0 / 3.0 : 0.75;
  cell_vjp_gather<LIM>(b, qs, lin, lout, (unsigned)c, gas, c_next, c0, stage == 3 ? 0 : (stage == 2 ? 1 : 2));
}

// ---------------------------------------------------------------------------
// host launcher
// ---------------------------------------------------------------------------
void launch_stage_vjp(const BlockDesc* blocks, const Tile* tiles, int ntiles, int stage, const Gas& gas, double dt,
                      int lim, hipStream_t st) {
  if (lim == LIM_VANLEER) {
    k_vjp_faces_visc<LIM_VANLEER><<<ntiles, 256, 0, st>>>(blocks, tiles, stage, gas, dt);
    k_vjp_cell<LIM_VANLEER><<<ntiles, 256, 0, st>>>(blocks, tiles, stage, gas);
  } else {
    k_vjp_faces_visc<LIM_MINMOD><<<ntiles, 256, 0, st>>>(blocks, tiles, stage, gas, dt);
    k_vjp_cell<LIM_MINMOD><<<ntiles, 256, 0, st>>>(blocks, tiles, stage, gas);
  }
}

}  // namespace famrdp
