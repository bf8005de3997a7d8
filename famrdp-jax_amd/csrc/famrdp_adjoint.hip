// Adjoint (reverse-mode) kernels of the TVD-RK3 step.
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
// The halo copy and the BCs are reversed afterwards (BC faces in reverse Face
// order, the halo through a source-sorted table).  Functions cite the forward
// code they reverse in famrdp_jax (rhs/invscd.py, rhs/viscous.py,
// physics/bc/*.py, mesh/halo.py, solver/step.py).
#include "famrdp_adjoint.h"

namespace famrdp {

// ---------------------------------------------------------------------------
// inviscid faces (rhs/invscd.py:53-131 reversed)
// ---------------------------------------------------------------------------
template <int LIM>
__device__ __forceinline__ void cell_vjp_faces(const BlockDesc& b, const double* __restrict__ qs,
                                               const double* __restrict__ lin, double c_mu, int64_t a, double g) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const int k = (int)(a % nk), j = (int)((a / nk) % nj), i = (int)(a / ((int64_t)nk * nj));
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double* out = b.fs + (int64_t)(d * 10) * N + a;
    const int ad = d == 0 ? i : (d == 1 ? j : k);
    const int nd = d == 0 ? ni : (d == 1 ? nj : nk);
    const int64_t sd = d == 0 ? (int64_t)nj * nk : (d == 1 ? (int64_t)nk : (int64_t)1);
    double lpl[5] = {0, 0, 0, 0, 0}, lpr[5] = {0, 0, 0, 0, 0};
    const bool va = ad >= 2 && ad <= nd - 3, vb = ad + 1 >= 2 && ad + 1 <= nd - 3;
    if (ad >= 1 && ad <= nd - 3 && (va || vb)) {
      const int64_t a1 = a + sd;
      double lFa[5];
      const double ra = va ? rcp_hd(b.vol[a] + 1e-300) : 0.0, rb = vb ? rcp_hd(b.vol[a1] + 1e-300) : 0.0;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double ma = va ? c_mu * lin[v * N + a] : 0.0, mb = vb ? c_mu * lin[v * N + a1] : 0.0;
        lFa[v] = -ra * ma + rb * mb;
      }
      double pv[4][5];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        double qq[5];
        const int64_t cc = a + (s - 1) * sd;
#pragma unroll
        for (int v = 0; v < 5; ++v) qq[v] = qs[v * N + cc];
        cons_to_prim_fast(qq, g, pv[s]);
      }
      double pL[5], pR[5];
      face_states<LIM>(pv[0], pv[1], pv[2], pv[3], pL, pR);
      double k0[3], k1[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) { k0[m] = b.kxyz[(d * 3 + m) * N + a]; k1[m] = b.kxyz[(d * 3 + m) * N + a1]; }
      double nu[3], nm;
      face_normal(k0, k1, nu, nm);
      double lF[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) lF[v] = lFa[v] * nm;
      roe_flux_prim_vjp(pL, pR, nu, g, lF, lpl, lpr);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) { out[v * N] = lpl[v]; out[(5 + v) * N] = lpr[v]; }
  }
}

// ---------------------------------------------------------------------------
// viscous flux tensor (rhs/viscous.py:86-108 and 30-84 reversed)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cell_vjp_visc(const BlockDesc& b, const double* __restrict__ qs,
                                              const double* __restrict__ lin, double c_mu, int64_t x,
                                              const Gas& gas) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const int k = (int)(x % nk), j = (int)((x / nk) % nj), i = (int)(x / ((int64_t)nk * nj));
  const int64_t sj = nk, si = (int64_t)nj * nk;
  const int64_t cm[3] = {x + (wrap_m(i, ni) - i) * si, x + (wrap_m(j, nj) - j) * sj, x + (wrap_m(k, nk) - k)};
  const int64_t cp[3] = {x + (wrap_p(i, ni) - i) * si, x + (wrap_p(j, nj) - j) * sj, x + (wrap_p(k, nk) - k)};
  // cotangent of G(x): the divergence at x - e_l uses +G(x)/2, at x + e_l uses -G(x)/2
  double lG[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int64_t ym = cm[l], yp = cp[l];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double mm = c_mu * lin[(r + 1) * N + ym], mp = c_mu * lin[(r + 1) * N + yp];
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        lG[gidx(r, jj)] += 0.5 * (b.kxyz[(l * 3 + jj) * N + ym] * mm - b.kxyz[(l * 3 + jj) * N + yp] * mp);
    }
  }
  double pv[5], d[3][4];
  load_prim_fast(qs, N, x, gas.gamma, pv);
  const double T = temperature_fast(pv, gas.R);
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    double pp[5], pm[5];
    load_prim_fast(qs, N, cp[l], gas.gamma, pp);
    load_prim_fast(qs, N, cm[l], gas.gamma, pm);
    d[l][0] = 0.5 * (pp[1] - pm[1]);
    d[l][1] = 0.5 * (pp[2] - pm[2]);
    d[l][2] = 0.5 * (pp[3] - pm[3]);
    d[l][3] = 0.5 * (temperature_fast(pp, gas.R) - temperature_fast(pm, gas.R));
  }
  double kk[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kk[m] = b.kxyz[m * N + x];
  double lu = 0, lv = 0, lw = 0, lT = 0, ld[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  viscous_flux_vjp(pv, T, d, kk, gas.gamma, gas.R, b.mu0, gas.T0, gas.Ts, gas.Pr_lam, lG, lu, lv, lw, lT, ld);
  double* out = b.vs + x;
  out[0] = lu; out[N] = lv; out[2 * N] = lw; out[3 * N] = lT;
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int s = 0; s < 4; ++s) out[(4 + 4 * l + s) * N] = ld[l][s];
}

// ---------------------------------------------------------------------------
// per-cell gather, primitive -> conservative, RK coefficients
//   mode 0: l0 = c0 lin,  lout = c_next lin + g     (stage 3)
//   mode 1: l0 += c0 lin, lout = c_next lin + g     (stage 2)
//   mode 2: l0 += lin + g                           (stage 1)
// ---------------------------------------------------------------------------
template <int LIM>
__device__ __forceinline__ void cell_vjp_gather(const BlockDesc& b, const double* __restrict__ qs,
                                                const double* lin, double* lout, int64_t y, const Gas& gas,
                                                double c_next, double c0, int mode) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const int k = (int)(y % nk), j = (int)((y / nk) % nj), i = (int)(y / ((int64_t)nk * nj));
  double lpv[5] = {0, 0, 0, 0, 0};
  // ---- inviscid: face states and MUSCL slopes along each axis
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double* fsd = b.fs + (int64_t)(d * 10) * N;
    const int yd = d == 0 ? i : (d == 1 ? j : k);
    const int nd = d == 0 ? ni : (d == 1 ? nj : nk);
    const int64_t sd = d == 0 ? (int64_t)nj * nk : (d == 1 ? (int64_t)nk : (int64_t)1);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      lpv[v] += fsd[v * N + y];                          // left state of face (y, y+1)
      if (yd >= 1) lpv[v] += fsd[(5 + v) * N + y - sd];  // right state of face (y-1, y)
    }
    // slopes s(x) = lim(pv(x)-pv(x-1), pv(x+1)-pv(x)) for x = y-1, y, y+1
    double pv[5][5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      const int xd = yd + s - 2;
      if (xd >= 0 && xd < nd) load_prim_fast(qs, N, y + (s - 2) * sd, gas.gamma, pv[s]);
      else
#pragma unroll
        for (int v = 0; v < 5; ++v) pv[s][v] = 0.0;
    }
#pragma unroll
    for (int o = -1; o <= 1; ++o) {
      const int xd = yd + o;
      if (xd < 1 || xd > nd - 2) continue;
      const int64_t xc = y + o * sd;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double ls = 0.5 * fsd[v * N + xc] - 0.5 * fsd[(5 + v) * N + xc - sd];
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
  const int64_t sj = nk, si = (int64_t)nj * nk;
  const int64_t cm[3] = {y + (wrap_m(i, ni) - i) * si, y + (wrap_m(j, nj) - j) * sj, y + (wrap_m(k, nk) - k)};
  const int64_t cp[3] = {y + (wrap_p(i, ni) - i) * si, y + (wrap_p(j, nj) - j) * sj, y + (wrap_p(k, nk) - k)};
  double ls4[4] = {b.vs[y], b.vs[N + y], b.vs[2 * N + y], b.vs[3 * N + y]};
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int s = 0; s < 4; ++s)
      ls4[s] += 0.5 * b.vs[(4 + 4 * l + s) * N + cm[l]] - 0.5 * b.vs[(4 + 4 * l + s) * N + cp[l]];
  double qy[5], pvy[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) qy[v] = qs[v * N + y];
  cons_to_prim_fast(qy, gas.gamma, pvy);
  lpv[1] += ls4[0]; lpv[2] += ls4[1]; lpv[3] += ls4[2];
  temperature_vjp(pvy, gas.R, ls4[3], lpv[0], lpv[4]);
  double gq[5] = {0, 0, 0, 0, 0};
  cons_to_prim_vjp(qy, gas.gamma, lpv, gq);
  // ---- RK coefficients (solver/step.py:45-53 reversed)
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double li = lin[v * N + y];
    if (mode == 2) {
      b.l0[v * N + y] += li + gq[v];
    } else {
      b.l0[v * N + y] = (mode == 0 ? 0.0 : b.l0[v * N + y]) + c0 * li;
      lout[v * N + y] = c_next * li + gq[v];
    }
  }
}

// Stage selector for the reverse sweep: which buffers a stage reads/writes.
//   stage 3: qs = q2, lin = la (cotangent of q_new), lout = lb
//   stage 2: qs = q1, lin = lb, lout = la
//   stage 1: qs = qh, lin = la
__device__ __forceinline__ void vjp_bufs(const BlockDesc& b, int stage, const double*& qs, const double*& lin,
                                         double*& lout) {
  qs = stage == 3 ? b.q2 : (stage == 2 ? b.q1 : b.qh);
  lin = stage == 3 ? b.la : (stage == 2 ? b.lb : b.la);
  lout = stage == 3 ? b.lb : (stage == 2 ? b.la : nullptr);
}

__device__ __forceinline__ double stage_cmu(int stage, double dt) {
  return stage == 3 ? (2.0 / 3.0) * dt : (stage == 2 ? 0.25 * dt : dt);
}

// The viscous part first: the face part reuses the neighbours' state and
// metrics it loaded (from registers or cache).  The scheduling barrier keeps
// the two parts' instructions apart (two waves per SIMD, as the face part
// alone needs; 4 registers spilled).
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
  cell_vjp_visc(b, qs, lin, cmu, c, gas);
  __builtin_amdgcn_sched_barrier(0);
  cell_vjp_faces<LIM>(b, qs, lin, cmu, c, gas.gamma);
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
  cell_vjp_gather<LIM>(b, qs, lin, lout, c, gas, c_next, c0, stage == 3 ? 0 : (stage == 2 ? 1 : 2));
}

// ---------------------------------------------------------------------------
// BC faces reversed (physics/bc/*.py): the ghost cotangent moves to the inner
// cell it was built from (or to q_inf for farfield) and is cleared.
// ---------------------------------------------------------------------------

// Copy of the inner layers of isothermal faces just before the face is
// applied (the reverse sweep linearises at those values).
__global__ void k_bc_snapshot(const BlockDesc* __restrict__ blocks, const BcEntry* __restrict__ entries, int ghost,
                              double* __restrict__ snap, int64_t maxcells) {
  const BcEntry e = entries[blockIdx.y];
  if (e.kind != BC_WALL_ISOTHERMAL) return;
  const BlockDesc b = blocks[e.block];
  const int ax = e.face / 2;
  const bool lo = (e.face % 2) == 0;
  const int n[3] = {b.ni, b.nj, b.nk};
  const int64_t st[3] = {(int64_t)b.nj * b.nk, b.nk, 1};
  const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const int64_t ncell = (int64_t)n[t1] * n[t2];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int64_t base = (t / n[t2]) * st[t1] + (t % n[t2]) * st[t2];
  for (int gl = 0; gl < ghost; ++gl) {
    const int inner = lo ? 2 * ghost - 1 - gl : n[ax] - 2 * ghost + gl;
    const int64_t ci = base + inner * st[ax];
    for (int v = 0; v < 5; ++v) snap[(((int64_t)blockIdx.y * ghost + gl) * 5 + v) * maxcells + t] = b.qh[v * b.N + ci];
  }
}

__global__ void k_bc_face_vjp(const BlockDesc* __restrict__ blocks, const BcEntry* __restrict__ entries, int ghost,
                              double gamma, double R, const double* __restrict__ snap, double* __restrict__ part,
                              int64_t maxcells) {
  const BcEntry e = entries[blockIdx.y];
  const BlockDesc b = blocks[e.block];
  const int ax = e.face / 2;
  const bool lo = (e.face % 2) == 0;
  const int n[3] = {b.ni, b.nj, b.nk};
  const int64_t st[3] = {(int64_t)b.nj * b.nk, b.nk, 1};
  const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const int64_t ncell = (int64_t)n[t1] * n[t2];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int64_t base = (t / n[t2]) * st[t1] + (t % n[t2]) * st[t2];
  const int na = n[ax];
  double* L = b.l0;
  const int64_t N = b.N;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int gl = 0; gl < ghost; ++gl) {
    const int outer = lo ? gl : na - 1 - gl;
    int inner = lo ? 2 * ghost - 1 - gl : na - 2 * ghost + gl;
    if (e.kind == BC_EXTRAPOLATE) inner = lo ? ghost : na - 1 - ghost;
    const int64_t co = base + outer * st[ax], ci = base + inner * st[ax];
    double lo5[5];
    for (int v = 0; v < 5; ++v) { lo5[v] = L[v * N + co]; L[v * N + co] = 0.0; }
    switch (e.kind) {
      case BC_WALL_ADIABATIC:
        for (int v = 0; v < 5; ++v) L[v * N + ci] += (v >= 1 && v <= 3) ? -lo5[v] : lo5[v];
        break;
      case BC_MIRROR_NORMAL:
        for (int v = 0; v < 5; ++v) L[v * N + ci] += (v == ax + 1) ? -lo5[v] : lo5[v];
        break;
      case BC_FARFIELD:
        for (int v = 0; v < 5; ++v) acc[v] += lo5[v];
        break;
      case BC_WALL_ISOTHERMAL: {
        double qi[5], li[5] = {0, 0, 0, 0, 0};
        for (int v = 0; v < 5; ++v) qi[v] = snap[(((int64_t)blockIdx.y * ghost + gl) * 5 + v) * maxcells + t];
        isothermal_vjp(qi, gamma, R, e.T_wall, lo5, li);
        for (int v = 0; v < 5; ++v) L[v * N + ci] += li[v];
        break;
      }
      default:  // BC_EXTRAPOLATE
        for (int v = 0; v < 5; ++v) L[v * N + ci] += lo5[v];
        break;
    }
  }
  if (e.kind == BC_FARFIELD)
    for (int v = 0; v < 5; ++v) part[((int64_t)blockIdx.y * 5 + v) * maxcells + t] = acc[v];
}

// Deterministic sum of the farfield partials of one face entry per workgroup,
// added to dqinf[(block*6 + face)*5 + v].
__global__ void __launch_bounds__(256) k_bc_reduce(const BlockDesc* __restrict__ blocks,
                                                   const BcEntry* __restrict__ entries, const double* __restrict__ part,
                                                   int64_t maxcells, double* __restrict__ dqinf) {
  const BcEntry e = entries[blockIdx.x];
  if (e.kind != BC_FARFIELD) return;
  const BlockDesc b = blocks[e.block];
  const int ax = e.face / 2;
  const int n[3] = {b.ni, b.nj, b.nk};
  const int64_t ncell = (int64_t)n[ax == 0 ? 1 : 0] * n[ax == 2 ? 1 : 2];
  __shared__ double red[5][256];
  for (int v = 0; v < 5; ++v) {
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < ncell; t += blockDim.x) s += part[((int64_t)blockIdx.x * 5 + v) * maxcells + t];
    red[v][threadIdx.x] = s;
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int v = 0; v < 5; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && dqinf)
    for (int v = 0; v < 5; ++v) dqinf[(e.block * 6 + e.face) * 5 + v] += red[v][0];
}

// ---------------------------------------------------------------------------
// halo copy reversed (mesh/halo.py:56-92): every source cell collects the
// cotangents of the ghost cells copied from it (fixed order), the ghost cells
// are cleared, then the sums are added to the sources.
// ---------------------------------------------------------------------------
struct HaloSrc {
  int32_t sb, sc, start, count;
};

__global__ void k_halo_vjp_gather(const BlockDesc* __restrict__ blocks, const HaloSrc* __restrict__ us, int nus,
                                  const HaloEntry* __restrict__ sorted, double* __restrict__ S) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nus) return;
  const HaloSrc h = us[u];
  double acc[5] = {0, 0, 0, 0, 0};
  for (int e = h.start; e < h.start + h.count; ++e) {
    const HaloEntry he = sorted[e];
    const BlockDesc b = blocks[he.tb];
    for (int v = 0; v < 5; ++v) acc[v] += b.l0[v * b.N + he.tc];
  }
  for (int v = 0; v < 5; ++v) S[(int64_t)v * nus + u] = acc[v];
}

__global__ void k_halo_vjp_clear(const BlockDesc* __restrict__ blocks, const HaloEntry* __restrict__ e, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const HaloEntry h = e[t];
  const BlockDesc b = blocks[h.tb];
  for (int v = 0; v < 5; ++v) b.l0[v * b.N + h.tc] = 0.0;
}

__global__ void k_halo_vjp_apply(const BlockDesc* __restrict__ blocks, const HaloSrc* __restrict__ us, int nus,
                                 const double* __restrict__ S) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nus) return;
  const HaloSrc h = us[u];
  const BlockDesc b = blocks[h.sb];
  for (int v = 0; v < 5; ++v) b.l0[v * b.N + h.sc] += S[(int64_t)v * nus + u];
}

// ---------------------------------------------------------------------------
// objectives (build-defined; the reference has none) and their gradients
//   kind 0: kinetic energy  J = sum_interior 1/2 |rho u|^2 / rho * vol
//   kind 1: wall pressure   J = sum_{i,j interior, k = ghost} p |grad zeta| vol
// grad (5, N) is written on every cell (zero where J does not depend on q).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_objective(int kind, const double* __restrict__ q,
                                                   const double* __restrict__ kxyz, const double* __restrict__ vol,
                                                   int ni, int nj, int nk, int ghost, double gamma,
                                                   double* __restrict__ grad, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t N = (int64_t)ni * nj * nk;
  double acc = 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(c % nk), j = (int)((c / nk) % nj), i = (int)(c / ((int64_t)nk * nj));
    const bool inij = i >= ghost && i < ni - ghost && j >= ghost && j < nj - ghost;
    const bool use = kind == 0 ? (inij && k >= ghost && k < nk - ghost) : (inij && k == ghost);
    double gq[5] = {0, 0, 0, 0, 0};
    if (use) {
      const double rho = q[c], m1 = q[N + c], m2 = q[2 * N + c], m3 = q[3 * N + c], E = q[4 * N + c];
      const double u = m1 / rho, v = m2 / rho, w = m3 / rho;
      const double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
      if (kind == 0) {
        const double V = vol[c];
        acc += ke * V;
        gq[0] = -0.5 * (u * u + v * v + w * w) * V; gq[1] = u * V; gq[2] = v * V; gq[3] = w * V;
      } else {
        const double k0 = kxyz[6 * N + c], k1 = kxyz[7 * N + c], k2 = kxyz[8 * N + c];
        const double A = sqrt(k0 * k0 + k1 * k1 + k2 * k2) * vol[c];
        const double p = (gamma - 1.0) * (E - ke);
        acc += p * A;
        const double s = (gamma - 1.0) * A;
        gq[0] = s * 0.5 * (u * u + v * v + w * w); gq[1] = -s * u; gq[2] = -s * v; gq[3] = -s * w; gq[4] = s;
      }
    }
    for (int vv = 0; vv < 5; ++vv) grad[vv * N + c] = gq[vv];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) s += part[t];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// Up to COPY_JOBS copies in one launch (passed by value as the kernel
// argument): job blockIdx.y copies n[job] doubles; 4 per thread.
constexpr int COPY_JOBS = 64;
struct CopyJobs {
  double* dst[COPY_JOBS];
  const double* src[COPY_JOBS];
  int64_t n[COPY_JOBS];
};

__global__ void __launch_bounds__(256) k_copy_jobs(CopyJobs jobs) {
  const int j = blockIdx.y;
  const int64_t n = jobs.n[j];
  double* __restrict__ d = jobs.dst[j];
  const double* __restrict__ s = jobs.src[j];
  for (int64_t t = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * 4; t < n; t += (int64_t)gridDim.x * 1024) {
    if (t + 3 < n) {
      const double a = s[t], b = s[t + 1], c = s[t + 2], e = s[t + 3];
      d[t] = a; d[t + 1] = b; d[t + 2] = c; d[t + 3] = e;
    } else {
      for (int64_t u = t; u < n; ++u) d[u] = s[u];
    }
  }
}

}  // namespace famrdp
