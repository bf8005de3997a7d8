// FP64 kernels of famrdp_jax's RHS + TVD-RK3 path for gfx950.
//
// Kernel families:
//   * unit kernels behind the reference's pointwise API (eos, MUSCL-2, Roe);
//   * the direct stage pair (viscous flux tensor, then inviscid + viscous
//     divergence + RK3 combination), one thread per padded cell, in a
//     per-block form and a batched form driven by a tile table;
//   * halo gather/scatter (pre-exchange semantics) and ghost-cell BC faces;
//   * grid metrics and the CFL / positivity reduction.
#include "famrdp_kernels.h"
#include "famrdp_tiles.h"

namespace famrdp {

// ---------------------------------------------------------------------------
// Unit kernels
// ---------------------------------------------------------------------------

__global__ void k_cons_to_prim(const double* __restrict__ q, double* __restrict__ pv, int64_t n, double g) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a[5], b[5];
  for (int v = 0; v < 5; ++v) a[v] = q[v * n + t];
  cons_to_prim(a, g, b);
  for (int v = 0; v < 5; ++v) pv[v * n + t] = b[v];
}

__global__ void k_prim_to_cons(const double* __restrict__ pv, double* __restrict__ q, int64_t n, double g) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a[5], b[5];
  for (int v = 0; v < 5; ++v) a[v] = pv[v * n + t];
  prim_to_cons(a, g, b);
  for (int v = 0; v < 5; ++v) q[v * n + t] = b[v];
}

__global__ void k_sutherland(const double* __restrict__ T, double* __restrict__ mu, int64_t n, double mu0,
                             double T0, double Ts) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) mu[t] = sutherland(T[t], mu0, T0, Ts);
}

__global__ void k_speed_of_sound(const double* __restrict__ rho, const double* __restrict__ p,
                                 double* __restrict__ c, int64_t n, double g) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) c[t] = sqrt(g * p[t] / rho[t]);
}

__global__ void k_limiter(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
                          int64_t n, int lim) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) out[t] = limit(lim, a[t], b[t]);
}

// pv (nvar, n0, n1, n2); faces along `axis` (1..3) -> (nvar, ..., n_axis-3, ...)
__global__ void k_muscl2(const double* __restrict__ pv, double* __restrict__ pL, double* __restrict__ pR,
                         int nvar, int n0, int n1, int n2, int axis, int lim) {
  const int f0 = axis == 1 ? n0 - 3 : n0, f1 = axis == 2 ? n1 - 3 : n1, f2 = axis == 3 ? n2 - 3 : n2;
  const int64_t nf = (int64_t)f0 * f1 * f2;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nf) return;
  const int a2 = (int)(t % f2), a1 = (int)((t / f2) % f1), a0 = (int)(t / ((int64_t)f2 * f1));
  const int64_t s0 = (int64_t)n1 * n2, s1 = n2;
  const int64_t stride = axis == 1 ? s0 : (axis == 2 ? s1 : 1);
  // face f along the axis sits between cells f+1 and f+2: base cell index = f
  const int64_t base = (int64_t)a0 * s0 + (int64_t)a1 * s1 + a2;
  const int64_t N = (int64_t)n0 * n1 * n2;
  for (int v = 0; v < nvar; ++v) {
    const double* p = pv + v * N + base;
    const double pmm = p[0], pc = p[stride], pp1 = p[2 * stride], pp2 = p[3 * stride];
    const double sL = limit(lim, pc - pmm, pp1 - pc);
    const double sR = limit(lim, pp1 - pc, pp2 - pp1);
    pL[v * nf + t] = pc + 0.5 * sL;
    pR[v * nf + t] = pp1 - 0.5 * sR;
  }
}

__global__ void k_roe_flux(const double* __restrict__ UL, const double* __restrict__ UR,
                           const double* __restrict__ nx, double* __restrict__ F, int64_t n, double g,
                           double eps_entropy) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a[5], b[5], nn[3], f[5];
  for (int v = 0; v < 5; ++v) { a[v] = UL[v * n + t]; b[v] = UR[v * n + t]; }
  for (int v = 0; v < 3; ++v) nn[v] = nx[v * n + t];
  roe_flux(a, b, nn, g, eps_entropy, f);
  for (int v = 0; v < 5; ++v) F[v * n + t] = f[v];
}

// ---------------------------------------------------------------------------
// Direct stage: per-cell device functions
// ---------------------------------------------------------------------------

// Viscous flux tensor G at cell c (rhs/viscous.py:30-108, roll wrap).
// FAST selects the stage-kernel arithmetic (reciprocals, fast sqrt).
template <bool FAST>
__device__ __forceinline__ void cell_viscous_flux(const double* __restrict__ q, const double* __restrict__ kxyz,
                                                  double* __restrict__ G, int ni, int nj, int nk, int64_t N,
                                                  int64_t c, const Gas& gas) {
  const int k = (int)(c % nk), j = (int)((c / nk) % nj), i = (int)(c / ((int64_t)nk * nj));
  const int64_t sj = nk, si = (int64_t)nj * nk;
  const int64_t cm[3] = {c + (wrap_m(i, ni) - i) * si, c + (wrap_m(j, nj) - j) * sj, c + (wrap_m(k, nk) - k)};
  const int64_t cp[3] = {c + (wrap_p(i, ni) - i) * si, c + (wrap_p(j, nj) - j) * sj, c + (wrap_p(k, nk) - k)};
  double pv[5];
  if (FAST) load_prim_fast(q, N, c, gas.gamma, pv); else load_prim(q, N, c, gas.gamma, pv);
  const double T = FAST ? temperature_fast(pv, gas.R) : temperature(pv, gas.R);
  double d[3][4];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    double pp[5], pm[5];
    if (FAST) {
      load_prim_fast(q, N, cp[l], gas.gamma, pp);
      load_prim_fast(q, N, cm[l], gas.gamma, pm);
    } else {
      load_prim(q, N, cp[l], gas.gamma, pp);
      load_prim(q, N, cm[l], gas.gamma, pm);
    }
    d[l][0] = 0.5 * (pp[1] - pm[1]);
    d[l][1] = 0.5 * (pp[2] - pm[2]);
    d[l][2] = 0.5 * (pp[3] - pm[3]);
    d[l][3] = FAST ? 0.5 * (temperature_fast(pp, gas.R) - temperature_fast(pm, gas.R))
                   : 0.5 * (temperature(pp, gas.R) - temperature(pm, gas.R));
  }
  double kk[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kk[m] = kxyz[m * N + c];
  double g9[9];
  if (FAST) viscous_flux_fast(pv, T, d, kk, gas, g9); else viscous_flux(pv, T, d, kk, gas, g9);
#pragma unroll
  for (int m = 0; m < 9; ++m) G[m * N + c] = g9[m];
}

// Inviscid RHS at cell c (rhs/invscd.py:53-131): axis contributions added in
// axis order, only where 2 <= c_d <= n_d - 3.
__device__ __forceinline__ void cell_invscd(const double* __restrict__ q, const double* __restrict__ kxyz,
                                            const double* __restrict__ vol, int ni, int nj, int nk, int64_t N,
                                            int64_t c, double g, int lim, double inv[5]) {
  const int k = (int)(c % nk), j = (int)((c / nk) % nj), i = (int)(c / ((int64_t)nk * nj));
  const int idx[3] = {i, j, k};
  const int n[3] = {ni, nj, nk};
  const int64_t st[3] = {(int64_t)nj * nk, nk, 1};
#pragma unroll
  for (int v = 0; v < 5; ++v) inv[v] = 0.0;
  const double rvol = vol[c] + 1e-300;
  for (int d = 0; d < 3; ++d) {
    if (idx[d] < 2 || idx[d] > n[d] - 3) continue;
    double pv[5][5];
#pragma unroll
    for (int s = 0; s < 5; ++s) load_prim(q, N, c + (s - 2) * st[d], g, pv[s]);
    double kd[3][3];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int m = 0; m < 3; ++m) kd[s][m] = kxyz[(d * 3 + m) * N + c + (s - 1) * st[d]];
    double Fm[5], Fp[5];
    face_flux(pv[0], pv[1], pv[2], pv[3], kd[0], kd[1], g, lim, Fm);   // face c-1/2
    face_flux(pv[1], pv[2], pv[3], pv[4], kd[1], kd[2], g, lim, Fp);   // face c+1/2
#pragma unroll
    for (int v = 0; v < 5; ++v) inv[v] = inv[v] + (-(Fp[v] - Fm[v]) / rvol);
  }
}

// Viscous divergence at cell c from G (rhs/viscous.py:86-108): for each row,
// sum over j of sum_l kxyz[l, j] * 0.5 (G_j(+e_l) - G_j(-e_l)).
__device__ __forceinline__ void cell_viscous_div(const double* __restrict__ G, const double* __restrict__ kxyz,
                                                 int ni, int nj, int nk, int64_t N, int64_t c, double vis[4]) {
  const int k = (int)(c % nk), j = (int)((c / nk) % nj), i = (int)(c / ((int64_t)nk * nj));
  const int64_t sj = nk, si = (int64_t)nj * nk;
  const int64_t cm[3] = {c + (wrap_m(i, ni) - i) * si, c + (wrap_m(j, nj) - j) * sj, c + (wrap_m(k, nk) - k)};
  const int64_t cp[3] = {c + (wrap_p(i, ni) - i) * si, c + (wrap_p(j, nj) - j) * sj, c + (wrap_p(k, nk) - k)};
  double kk[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kk[m] = kxyz[m * N + c];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double res = 0.0;
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      const double* Gc = G + gidx(r, jj) * N;
      const double d0 = 0.5 * (Gc[cp[0]] - Gc[cm[0]]);
      const double d1 = 0.5 * (Gc[cp[1]] - Gc[cm[1]]);
      const double d2 = 0.5 * (Gc[cp[2]] - Gc[cm[2]]);
      res = res + (d0 * kk[0 * 3 + jj] + d1 * kk[1 * 3 + jj] + d2 * kk[2 * 3 + jj]);
    }
    vis[r] = res;
  }
}

// Shu-Osher combination (solver/step.py:45-53).
__device__ __forceinline__ double rk_combine(int stage, double q0, double qs, double dt, double r) {
  if (stage == 1) return qs + dt * r;
  if (stage == 2) return 0.75 * q0 + 0.25 * (qs + dt * r);
  return (1.0 / 3.0) * q0 + (2.0 / 3.0) * (qs + dt * r);
}

// q0 may alias qout (stage 3 writes the state in place): both are read for
// cell c only, and every read happens before the first write.
__device__ __forceinline__ void cell_stage(const double* __restrict__ qin, const double* q0, double* qout,
                                           const double* __restrict__ kxyz, const double* __restrict__ vol,
                                           const double* __restrict__ G, int ni, int nj, int nk, int64_t N, int64_t c,
                                           double g, int lim, double dt, int stage) {
  double inv[5], vis[4];
  cell_invscd(qin, kxyz, vol, ni, nj, nk, N, c, g, lim, inv);
  cell_viscous_div(G, kxyz, ni, nj, nk, N, c, vis);
  double rhs[5];
  rhs[0] = inv[0] + 0.0;
  rhs[1] = inv[1] + vis[0];
  rhs[2] = inv[2] + vis[1];
  rhs[3] = inv[3] + vis[2];
  rhs[4] = inv[4] + vis[3];
  double qs[5], qq0[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    qs[v] = qin[v * N + c];
    qq0[v] = stage == 1 ? 0.0 : q0[v * N + c];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) qout[v * N + c] = rk_combine(stage, qq0[v], qs[v], dt, rhs[v]);
}

// ---------------------------------------------------------------------------
// Direct stage kernels: per-block launches
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_viscous_flux(const double* __restrict__ q, const double* __restrict__ kxyz,
                                                      double* __restrict__ G, int ni, int nj, int nk, Gas gas) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < N) cell_viscous_flux<false>(q, kxyz, G, ni, nj, nk, N, c, gas);
}

__global__ void __launch_bounds__(256) k_viscous_flux_fast(const double* __restrict__ q,
                                                           const double* __restrict__ kxyz, double* __restrict__ G,
                                                           int ni, int nj, int nk, Gas gas) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < N) cell_viscous_flux<true>(q, kxyz, G, ni, nj, nk, N, c, gas);
}

__global__ void k_fastmath_probe(const double* __restrict__ x, double* __restrict__ r, double* __restrict__ s,
                                 double* __restrict__ q, int64_t n) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) {
    r[t] = rcp(x[t]);
    s[t] = fsqrt(x[t]);
    if (q) q[t] = rsq_acc(x[t]);
  }
}

__global__ void k_fastmath_probe1(const double* __restrict__ x, double* __restrict__ r, double* __restrict__ q,
                                  int64_t n) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) {
    r[t] = rcp1(x[t]);
    q[t] = rsq1(x[t]);
  }
}

__global__ void __launch_bounds__(256) k_stage(const double* __restrict__ qin, const double* q0,
                                               double* qout, const double* __restrict__ kxyz,
                                               const double* __restrict__ vol, const double* __restrict__ G, int ni,
                                               int nj, int nk, double g, int lim, double dt, int stage) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < N) cell_stage(qin, q0, qout, kxyz, vol, G, ni, nj, nk, N, c, g, lim, dt, stage);
}

// RHS only (the unit entry points compute_invscd_rhs / compute_viscous_rhs)
__global__ void __launch_bounds__(256) k_invscd_rhs(const double* __restrict__ q, const double* __restrict__ kxyz,
                                                    const double* __restrict__ vol, double* __restrict__ rhs, int ni,
                                                    int nj, int nk, double g, int lim) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= N) return;
  double inv[5];
  cell_invscd(q, kxyz, vol, ni, nj, nk, N, c, g, lim, inv);
  for (int v = 0; v < 5; ++v) rhs[v * N + c] = inv[v];
}

__global__ void __launch_bounds__(256) k_viscous_div(const double* __restrict__ G, const double* __restrict__ kxyz,
                                                     double* __restrict__ rhs, int ni, int nj, int nk) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= N) return;
  double vis[4];
  cell_viscous_div(G, kxyz, ni, nj, nk, N, c, vis);
  rhs[c] = 0.0;
  for (int v = 0; v < 4; ++v) rhs[(v + 1) * N + c] = vis[v];
}

// Generic RK combination for rk3_step(q, L, dt): out = a*x + b*(y + c*z), or
// y + c*z when mode == 1 (time/rk.py:17-19).
__global__ void k_axpby(double* __restrict__ out, double a, const double* __restrict__ x, double b,
                        const double* __restrict__ y, double c, const double* __restrict__ z, int64_t n, int mode) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = mode == 1 ? y[t] + c * z[t] : a * x[t] + b * (y[t] + c * z[t]);
}

// ---------------------------------------------------------------------------
// Batched direct stage kernels: persistent grid over a tile table that spans
// every block of the plan (the block-offset table is `blocks`).
// ---------------------------------------------------------------------------

template <bool FAST>
__global__ void __launch_bounds__(256) k_viscous_flux_batched(const BlockDesc* __restrict__ blocks,
                                                              const Tile* __restrict__ tiles, int ntiles, int stage,
                                                              Gas gas) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const BlockDesc& b = blocks[tl.block];
    const int64_t c = (int64_t)tl.start + threadIdx.x;
    if (c < b.N) {
      const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
      Gas gb = gas;
      gb.mu0 = b.mu0;
      cell_viscous_flux<FAST>(qin, b.kxyz, b.G, b.ni, b.nj, b.nk, b.N, c, gb);
    }
  }
}

__global__ void __launch_bounds__(256) k_stage_batched(const BlockDesc* __restrict__ blocks,
                                                       const Tile* __restrict__ tiles, int ntiles, double g, int lim,
                                                       double dt, int stage) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const BlockDesc& b = blocks[tl.block];
    const int64_t c = (int64_t)tl.start + threadIdx.x;
    if (c < b.N) {
      const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
      double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
      cell_stage(qin, b.q, qout, b.kxyz, b.vol, b.G, b.ni, b.nj, b.nk, b.N, c, g, lim, dt * b.dt_scale, stage);
    }
  }
}

// ---------------------------------------------------------------------------
// Halo exchange (mesh/halo.py:56-92): gather every final ghost value from the
// pre-exchange state into `stage`, then scatter; two launches give the
// reference's "all copies read the original arrays" semantics.
// ---------------------------------------------------------------------------

__global__ void k_halo_gather(const BlockDesc* __restrict__ blocks, const HaloEntry* __restrict__ e, int n,
                              double* __restrict__ stage) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const HaloEntry h = e[t];
  const BlockDesc s = blocks[h.sb];
#pragma unroll
  for (int v = 0; v < 5; ++v) stage[(int64_t)v * n + t] = s.q[v * s.N + h.sc];
}

__global__ void k_halo_scatter(const BlockDesc* __restrict__ blocks, const HaloEntry* __restrict__ e, int n,
                               const double* __restrict__ stage) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const HaloEntry h = e[t];
  const BlockDesc b = blocks[h.tb];
#pragma unroll
  for (int v = 0; v < 5; ++v) b.q[v * b.N + h.tc] = stage[(int64_t)v * n + t];
}

// ---------------------------------------------------------------------------
// Ghost-cell BCs (physics/bc/*.py).  One launch per face index (Face order),
// covering that face of every block that has a BC there; blockIdx.y = entry.
// ---------------------------------------------------------------------------

__global__ void k_bc_face(const BlockDesc* __restrict__ blocks, const BcEntry* __restrict__ entries, int ghost,
                          double gamma, double R) {
  const BcEntry e = entries[blockIdx.y];
  const BlockDesc b = blocks[e.block];
  const int ax = e.face / 2;            // 0: i, 1: j, 2: k
  const bool lo = (e.face % 2) == 0;
  const int n[3] = {b.ni, b.nj, b.nk};
  const int64_t st[3] = {(int64_t)b.nj * b.nk, b.nk, 1};
  const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const int64_t ncell = (int64_t)n[t1] * n[t2];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int a = (int)(t / n[t2]), c2 = (int)(t % n[t2]);
  const int64_t base = a * st[t1] + c2 * st[t2];
  const int na = n[ax];
  for (int gl = 0; gl < ghost; ++gl) {
    // bc/base.py:21-33 mirror pairs; extrapolation reads the first interior layer
    const int outer = lo ? gl : na - 1 - gl;
    int inner = lo ? 2 * ghost - 1 - gl : na - 2 * ghost + gl;
    if (e.kind == BC_EXTRAPOLATE) inner = lo ? ghost : na - 1 - ghost;
    const int64_t co = base + outer * st[ax], ci = base + inner * st[ax];
    double qi[5], qo[5];
    for (int v = 0; v < 5; ++v) qi[v] = b.q[v * b.N + ci];
    switch (e.kind) {
      case BC_WALL_ADIABATIC:
        qo[0] = qi[0]; qo[1] = -qi[1]; qo[2] = -qi[2]; qo[3] = -qi[3]; qo[4] = qi[4];
        break;
      case BC_WALL_ISOTHERMAL: {   // bc/wall.py:35-58
        const double rho_i = qi[0];
        const double ke_i = 0.5 * (qi[1] * qi[1] + qi[2] * qi[2] + qi[3] * qi[3]) / rho_i;
        const double p_g = (gamma - 1.0) * (qi[4] - ke_i);
        const double rho_g = p_g / (R * e.T_wall);
        const double scale = rho_g / rho_i;
        const double mu = -qi[1] * scale, mv = -qi[2] * scale, mw = -qi[3] * scale;
        const double ke_g = 0.5 * (mu * mu + mv * mv + mw * mw) / rho_g;
        qo[0] = rho_g; qo[1] = mu; qo[2] = mv; qo[3] = mw; qo[4] = p_g / (gamma - 1.0) + ke_g;
        break;
      }
      case BC_MIRROR_NORMAL:
        for (int v = 0; v < 5; ++v) qo[v] = (v == ax + 1) ? qi[v] * -1.0 : qi[v] * 1.0;
        break;
      case BC_FARFIELD:
        for (int v = 0; v < 5; ++v) qo[v] = e.q_inf[v] * 1.0;
        break;
      default:  // BC_EXTRAPOLATE
        for (int v = 0; v < 5; ++v) qo[v] = qi[v];
        break;
    }
    for (int v = 0; v < 5; ++v) b.q[v * b.N + co] = qo[v];
  }
}

// ---------------------------------------------------------------------------
// Grid metrics (mesh/metric.py:18-70): 4th-order central differences with
// roll wrap, J[m][l] = dx_m/dxi_l, jac = det J, kxyz = inv(J) as [l][m].
// ---------------------------------------------------------------------------

__global__ void k_metrics(const double* __restrict__ xyz, double* __restrict__ jac, double* __restrict__ kxyz,
                          double* __restrict__ vol, int ni, int nj, int nk) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int k = (int)(c % nk), j = (int)((c / nk) % nj), i = (int)(c / ((int64_t)nk * nj));
  const int idx[3] = {i, j, k};
  const int n[3] = {ni, nj, nk};
  const int64_t st[3] = {(int64_t)nj * nk, nk, 1};
  double J[3][3];
  for (int l = 0; l < 3; ++l) {
    int64_t o[4];
    const int sh[4] = {2, 1, -1, -2};
    for (int s = 0; s < 4; ++s) {
      int ii = idx[l] + sh[s];
      ii = ((ii % n[l]) + n[l]) % n[l];
      o[s] = c + (int64_t)(ii - idx[l]) * st[l];
    }
    for (int m = 0; m < 3; ++m) {
      const double* f = xyz + m * N;
      J[m][l] = (-f[o[0]] + 8.0 * f[o[1]] - 8.0 * f[o[2]] + f[o[3]]) / 12.0;
    }
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double inv[3][3] = {
      {c00 / det, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det},
      {c01 / det, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det},
      {c02 / det, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det}};
  jac[c] = det;
  vol[c] = fabs(det);
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m) kxyz[(l * 3 + m) * N + c] = inv[l][m];
}

// ---------------------------------------------------------------------------
// CFL time step and positivity monitor over interior cells (two passes, no
// atomics): per-workgroup max spectral radius and bad-cell count, then one
// workgroup reduces them.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_cfl_partial(const double* __restrict__ q, const double* __restrict__ kxyz,
                                                     const double* __restrict__ vol, int ni, int nj, int nk, int ghost,
                                                     Gas gas, double* __restrict__ part_lam,
                                                     double* __restrict__ part_bad) {
  __shared__ double slam[256];
  __shared__ double sbad[256];
  const int64_t N = (int64_t)ni * nj * nk;
  const int ii = ni - 2 * ghost, jj = nj - 2 * ghost, kk = nk - 2 * ghost;
  const int64_t nint = (int64_t)ii * jj * kk;
  double lam = 0.0, bad = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nint; t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % kk) + ghost, j = (int)((t / kk) % jj) + ghost, i = (int)(t / ((int64_t)kk * jj)) + ghost;
    const int64_t c = ((int64_t)i * nj + j) * nk + k;
    double pv[5];
    load_prim(q, N, c, gas.gamma, pv);
    const bool ok = isfinite(pv[0]) && isfinite(pv[1]) && isfinite(pv[2]) && isfinite(pv[3]) && isfinite(pv[4]) &&
                    pv[0] > 0.0 && pv[4] > 0.0;
    if (!ok) { bad += 1.0; continue; }
    const double cs = sqrt(gas.gamma * pv[4] / pv[0]);
    const double T = pv[4] / (pv[0] * gas.R);
    const double mu = sutherland(T, gas.mu0, gas.T0, gas.Ts);
    double inv = 0.0, ksq = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double k0 = kxyz[(d * 3 + 0) * N + c], k1 = kxyz[(d * 3 + 1) * N + c], k2 = kxyz[(d * 3 + 2) * N + c];
      const double kn2 = k0 * k0 + k1 * k1 + k2 * k2;
      inv = inv + (fabs(pv[1] * k0 + pv[2] * k1 + pv[3] * k2) + cs * sqrt(kn2));
      ksq = ksq + kn2;
    }
    const double l = inv / vol[c] + fmax(4.0 / 3.0, gas.gamma / gas.Pr_lam) * mu / pv[0] * ksq;
    lam = fmax(lam, l);
  }
  slam[threadIdx.x] = lam;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      slam[threadIdx.x] = fmax(slam[threadIdx.x], slam[threadIdx.x + s]);
      sbad[threadIdx.x] += sbad[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_lam[blockIdx.x] = slam[0];
    part_bad[blockIdx.x] = sbad[0];
  }
}

__global__ void __launch_bounds__(256) k_cfl_final(const double* __restrict__ part_lam,
                                                   const double* __restrict__ part_bad, int n, double cfl,
                                                   double* __restrict__ out) {
  __shared__ double slam[256];
  __shared__ double sbad[256];
  double lam = 0.0, bad = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    lam = fmax(lam, part_lam[t]);
    bad += part_bad[t];
  }
  slam[threadIdx.x] = lam;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      slam[threadIdx.x] = fmax(slam[threadIdx.x], slam[threadIdx.x + s]);
      sbad[threadIdx.x] += sbad[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = cfl / slam[0];
    out[1] = slam[0];
    out[2] = sbad[0];
  }
}

// ---------------------------------------------------------------------------
// State monitor (Config.nan_check): interior cells of every block of a plan
// with a non-finite conservative value, rho <= 0 or p <= 0, counted per
// workgroup over the plan's tile table (no atomics), then summed.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_monitor_partial(const BlockDesc* __restrict__ blocks,
                                                         const Tile* __restrict__ tiles, int ntiles, int ghost,
                                                         double gamma, double* __restrict__ part) {
  __shared__ double sbad[256];
  double bad = 0.0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const BlockDesc& b = blocks[tl.block];
    const int64_t c = (int64_t)tl.start + threadIdx.x;
    if (c >= b.N) continue;
    const int k = (int)(c % b.nk), j = (int)((c / b.nk) % b.nj), i = (int)(c / ((int64_t)b.nk * b.nj));
    if (i < ghost || j < ghost || k < ghost || i >= b.ni - ghost || j >= b.nj - ghost || k >= b.nk - ghost) continue;
    double pv[5];
    load_prim(b.q, b.N, c, gamma, pv);
    const bool ok = isfinite(pv[0]) && isfinite(pv[1]) && isfinite(pv[2]) && isfinite(pv[3]) && isfinite(pv[4]) &&
                    pv[0] > 0.0 && pv[4] > 0.0;
    bad += ok ? 0.0 : 1.0;
  }
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sbad[threadIdx.x] += sbad[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sbad[0];
}

__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) a += part[t];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

}  // namespace famrdp

// ===========================================================================
// Fused stage kernel: one launch per RK3 stage, every face flux and every
// viscous flux tensor computed once per tile.
//
// A workgroup owns a TI x TJ x TK tile (one thread per cell).
//  1. Stage primitives in LDS for the tile, a full one-cell halo box and a
//     second halo layer along each axis.  Halo cells are read at wrapped
//     (periodic) block coordinates: that is the reference's jnp.roll for the
//     viscous terms, and the inviscid faces never reach a wrapped cell (they
//     exist only between cells 1..n-2 of their axis).
//  2. Viscous flux tensor G (rhs/viscous.py:30-108) for the tile plus a
//     one-cell cross halo into LDS; every cell takes its viscous divergence.
//  3. For each axis in the reference's order (i, j, k): MUSCL-2 + Roe face
//     fluxes of the tile's low faces plus one layer of high faces into an LDS
//     face buffer (aliasing G), then -(F(c+1/2) - F(c-1/2)) / vol for cells
//     with 2 <= c_d <= n_d - 3 (rhs/invscd.py:100-131).
//  4. rhs = inv + vis and the Shu-Osher combination (solver/step.py:45-53).
// ===========================================================================
namespace famrdp {

// ---------------------------------------------------------------------------
// Inviscid faces of one axis D for a tile (rhs/invscd.py:53-131), shared by
// the fused and split stage kernels.  Faces of the tile: the low face of
// every cell (computed by the cell's own thread, with kxyz of the cell from
// registers) and the layer of high faces at the far end of the tile (dealt
// to the threads beyond the tile's cells, or after the low face when there
// are none).  Face fluxes go to the LDS buffer s_f (axes i, j: f * Pd + o,
// axis k: o * (TK + 1) + f, face f between local cells f-1 and f); after a
// barrier every cell with 2 <= c <= n-3 along D takes -(F(c+1/2) -
// F(c-1/2)) / vol.  Ends with a barrier (s_f is reused by the next axis).
// ---------------------------------------------------------------------------
// kxyz[D, :] of the left neighbour of cell c along D (the left cell of the
// cell's own low face), loaded early so the face phase does not wait on it;
// zero where that face does not exist.
template <int D>
__device__ __forceinline__ void load_kl_own(const BlockDesc& b, int64_t c, int gl_d, bool own, double kl[3]) {
  const int64_t st = D == 0 ? (int64_t)b.nj * b.nk : (D == 1 ? (int64_t)b.nk : (int64_t)1);
  const bool ok = own && gl_d >= 1;
  const int64_t cl = ok ? c - st : c;   // always a valid address: a fixed number of loads
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double v = b.kxyz[(D * 3 + m) * b.N + cl];
    kl[m] = ok ? v : 0.0;
  }
}

template <int D, class T, int TI, int TJ, int TK, int NTH, int LIM>
__device__ __forceinline__ void face_axis(const BlockDesc& b, int i0, int j0, int k0, const double* s_pv,
                                          double* s_f, const double kk[9], const double kl_own[3], bool own, int t,
                                          int x, int y, int z, double rvol, double g, double inv[5]) {
  constexpr int NT = TI * TJ * TK;
  constexpr int Td = D == 0 ? TI : (D == 1 ? TJ : TK);
  constexpr int Pd = NT / Td;
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const int64_t si = (int64_t)nj * nk, sj = nk;
  const int n_d = D == 0 ? ni : (D == 1 ? nj : nk);
  const int org_d = D == 0 ? i0 : (D == 1 ? j0 : k0);
  constexpr int64_t one = 1;
  const int64_t st_d = D == 0 ? si : (D == 1 ? sj : one);
  auto face = [&](int f, int lx, int ly, int lz, int slot, bool kr_own) {
    const int ai = i0 + lx, aj = j0 + ly, ak = k0 + lz;
    const int a = org_d + f - 1;          // cell left of the face (unwrapped)
    if (!(ai < ni && aj < nj && ak < nk && a >= 1 && a <= n_d - 3)) return;
    double p4[4][5];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int sl = T::template axis_slot<D>(lx + (D == 0 ? s4 - 2 : 0), ly + (D == 1 ? s4 - 2 : 0),
                                              lz + (D == 2 ? s4 - 2 : 0));
#pragma unroll
      for (int v = 0; v < 5; ++v) p4[s4][v] = s_pv[v * T::NPVS + sl];
    }
    const int64_t cr = (int64_t)ai * si + (int64_t)aj * sj + ak;
    const int64_t cl = cr - st_d;    // a >= 1: in the block
    double kl[3], kr[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      kl[m] = kr_own ? kl_own[m] : b.kxyz[(D * 3 + m) * N + cl];
      kr[m] = kr_own ? kk[D * 3 + m] : b.kxyz[(D * 3 + m) * N + cr];
    }
    double Fa[5];
    face_flux_prim<LIM>(p4[0], p4[1], p4[2], p4[3], kl, kr, g, Fa);
#pragma unroll
    for (int v = 0; v < 5; ++v) s_f[v * T::NF + slot] = Fa[v];
  };
  const int loc_d = D == 0 ? x : (D == 1 ? y : z);
  const int o_own = D == 0 ? (y * TK + z) : (D == 1 ? (x * TK + z) : (x * TJ + y));
  if (t < NT) face(loc_d, x, y, z, D == 2 ? o_own * (TK + 1) + loc_d : loc_d * Pd + o_own, true);
  for (int h = (t >= NT ? t - NT : t + NTH - NT); h < Pd; h += NTH) {
    int lx, ly, lz;
    if constexpr (D == 0) { lx = TI; ly = h / TK; lz = h % TK; }
    else if constexpr (D == 1) { lx = h / TK; ly = TJ; lz = h % TK; }
    else { lx = h / TJ; ly = h % TJ; lz = TK; }
    face(Td, lx, ly, lz, D == 2 ? h * (TK + 1) + Td : Td * Pd + h, false);
  }
  __syncthreads();
  const int gl_d = D == 0 ? i0 + x : (D == 1 ? j0 + y : k0 + z);
  if (own && gl_d >= 2 && gl_d <= n_d - 3) {
    const int lo = D == 2 ? o_own * (TK + 1) + loc_d : loc_d * Pd + o_own;
    const int hi = D == 2 ? lo + 1 : lo + Pd;
#pragma unroll
    for (int v = 0; v < 5; ++v) inv[v] = inv[v] + (-(s_f[v * T::NF + hi] - s_f[v * T::NF + lo]) * rvol);
  }
  __syncthreads();
}

template <int TI, int TJ, int TK, int W, bool ALL>
__device__ __forceinline__ void fused_cells_fill(const BlockDesc& b, int i0, int j0, int k0, int t,
                                                 FusedCells<TI, TJ, TK>& fc) {
  using C = FusedCells<TI, TJ, TK>;
  using T = typename C::T;
  const int ni = b.ni, nj = b.nj, nk = b.nk;
#pragma unroll
  for (int it = 0; it < C::NIT; ++it) {
    const int h = t + it * T::NT;
    const Loc3 l = T::pv_cell(h < T::NPV ? h : T::NPV - 1);
    fc.stage[it] = tile_cell_rel<W>(l.x + 2, l.y + 2, l.z + 2, i0, j0, k0, ni, nj, nk);
  }
  if constexpr (!ALL) return;
#pragma unroll
  for (int it = 1; it < C::NGIT; ++it) {
    const int h = t + it * T::NT;
    const Loc3 l = T::g_cell(h < T::NG ? h : T::NG - 1);
    fc.gh[it - 1] = tile_cell_rel<W>(l.x + 2, l.y + 2, l.z + 2, i0, j0, k0, ni, nj, nk);
  }
  const int x = t / (TJ * TK), y = (t / TK) % TJ, z = t % TK;
  fc.c = tile_cell<W>(i0 + x, j0 + y, k0 + z, ni, nj, nk);
}

// ALL = false: the staging slots only, for thread id t
template <int TI, int TJ, int TK, bool ALL = true>
__device__ __forceinline__ void fused_cells(const BlockDesc& b, int i0, int j0, int k0, int t,
                                            FusedCells<TI, TJ, TK>& fc) {
  const int w = fused_wrap_mode<TI, TJ, TK>(b.wrap_mod, i0, j0, k0, b.ni, b.nj, b.nk);
  if (w == WRAP_MOD) fused_cells_fill<TI, TJ, TK, WRAP_MOD, ALL>(b, i0, j0, k0, t, fc);
  else if (w == WRAP_NONE) fused_cells_fill<TI, TJ, TK, WRAP_NONE, ALL>(b, i0, j0, k0, t, fc);
  else fused_cells_fill<TI, TJ, TK, WRAP_ONCE, ALL>(b, i0, j0, k0, t, fc);
}

// v, hidden from the optimiser: values computed from it are computed again
// rather than kept in registers from an earlier, identical computation
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Staging loads of one tile into registers.  Every thread issues the same
// number of loads, so the compiler can wait for older loads without waiting
// for these.
template <int TI, int TJ, int TK>
__device__ __forceinline__ void fused_stage_loads(const BlockDesc& b, const double* __restrict__ qin,
                                                  const FusedCells<TI, TJ, TK>& fc, double qq[][5]) {
#pragma unroll
  for (int it = 0; it < FusedCells<TI, TJ, TK>::NIT; ++it) {
#pragma unroll
    for (int v = 0; v < 5; ++v) qq[it][v] = qin[v * b.N + fc.stage[it]];
  }
}

// One RK stage of one tile whose staging values are in qq (fused_stage_loads
// of this tile).  Global loads are grouped by the phase that first needs
// them: metrics of the G cells before the G phase, the faces' neighbour
// metrics and volume before the faces, the state at the update.
// RS (re-staging): the viscous tensors are kept in registers until every
// thread is done reading the primitives, then written over them; after the
// divergence the primitives are staged again (their loads issued before the
// divergence) for the faces.  Less LDS per workgroup, so more share a CU.
template <int TI, int TJ, int TK, int LIM, bool RS>
__device__ __forceinline__ void stage_tile_fused_regs(const BlockDesc& b, int i0, int j0, int k0,
                                                      const double* __restrict__ qin, const double* q0, double* qout,
                                                      const Gas& gas, double dt, int stage,
                                                      const FusedCells<TI, TJ, TK>& fc, double qq[][5]) {
  using T = FusedGeom<TI, TJ, TK>;
  constexpr int NIT = FusedCells<TI, TJ, TK>::NIT;
  __shared__ double lds[RS ? T::NLDS_RS : T::NLDS];
  double* s_pv = lds;                                  // [5][NPV]
  double* s_T = lds + 5 * T::NPV;                      // [NPV]   temperature of the staged cells
  double* s_g = RS ? lds : lds + 6 * T::NPV;           // [9][NG] (phase 2)
  double* s_f = RS ? lds + 5 * T::NPV : s_g;           // [5][NF] (phase 3)
  const int t = threadIdx.x;
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const double g = gas.gamma;
  Gas gb = gas;
  gb.mu0 = b.mu0;
  dt = dt * b.dt_scale;
  const int x = t / (TJ * TK), y = (t / TK) % TJ, z = t % TK;
  const int gi = i0 + x, gj = j0 + y, gk = k0 + z;
  const bool own = gi < ni && gj < nj && gk < nk;
  // the tile cell of this thread (wrapped into the block when outside it)
  const int c = fc.c;

  // ---- 1. primitives from the staged registers --------------------------------
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int h = t + it * T::NT;
    if (h < T::NPV) {
      double pp[5];
      cons_to_prim_fast(qq[it], g, pp);
#pragma unroll
      for (int v = 0; v < 5; ++v) s_pv[v * T::NPV + h] = pp[v];
      s_T[h] = temperature_fast(pp, gb.R);
    }
  }

  // ---- metrics of the G cells --------------------------------------------------
  double kk[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kk[m] = b.kxyz[m * N + c];
  constexpr int NGIT = FusedCells<TI, TJ, TK>::NGIT;
  double kh[NGIT > 1 ? NGIT - 1 : 1][9];
#pragma unroll
  for (int it = 1; it < NGIT; ++it) {
#pragma unroll
    for (int m = 0; m < 9; ++m) kh[it - 1][m] = b.kxyz[m * N + fc.gh[it - 1]];
  }
  __syncthreads();

  // ---- 2. viscous flux tensor and its divergence -----------------------------
  double gsave[RS ? NGIT : 1][9];
#pragma unroll
  for (int it = 0; it < NGIT; ++it) {
    const int h = t + it * T::NT;
    if (h >= T::NG) break;
    const Loc3 lc = T::g_cell(h);
    const int lx = lc.x, ly = lc.y, lz = lc.z;
    double kc[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) kc[m] = it == 0 ? kk[m] : kh[it > 0 ? it - 1 : 0][m];
    double d[3][4];
    const int s0 = T::pv_slot(lx, ly, lz);
    const double pv[5] = {0.0, s_pv[T::NPV + s0], s_pv[2 * T::NPV + s0], s_pv[3 * T::NPV + s0], 0.0};
    const double Tc = s_T[s0];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const int sp = T::pv_slot(lx + (l == 0), ly + (l == 1), lz + (l == 2));
      const int sm = T::pv_slot(lx - (l == 0), ly - (l == 1), lz - (l == 2));
      d[l][0] = 0.5 * (s_pv[T::NPV + sp] - s_pv[T::NPV + sm]);
      d[l][1] = 0.5 * (s_pv[2 * T::NPV + sp] - s_pv[2 * T::NPV + sm]);
      d[l][2] = 0.5 * (s_pv[3 * T::NPV + sp] - s_pv[3 * T::NPV + sm]);
      d[l][3] = 0.5 * (s_T[sp] - s_T[sm]);
    }
    double g9[9];
    viscous_flux_fast(pv, Tc, d, kc, gb, g9);
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      if constexpr (RS) gsave[it][m] = g9[m];
      else s_g[m * T::NG + h] = g9[m];
    }
  }
  __syncthreads();
  if constexpr (RS) {   // the primitives are no longer read: G goes over them
#pragma unroll
    for (int it = 0; it < NGIT; ++it) {
      const int h = t + it * T::NT;
      if (h < T::NG) {
#pragma unroll
        for (int m = 0; m < 9; ++m) s_g[m * T::NG + h] = gsave[it][m];
      }
    }
    // indices of the second staging computed again (held through the G
    // phase they would cost registers the three-per-CU budget lacks)
    FusedCells<TI, TJ, TK> fc2;
    fused_cells<TI, TJ, TK, false>(b, i0, j0, k0, opaque(t), fc2);
    fused_stage_loads<TI, TJ, TK>(b, qin, fc2, qq);
    __syncthreads();
  }
  double vis[4] = {0.0, 0.0, 0.0, 0.0};
  if (own) {
    const int sp[3] = {T::g_slot(x + 1, y, z), T::g_slot(x, y + 1, z), T::g_slot(x, y, z + 1)};
    const int sm[3] = {T::g_slot(x - 1, y, z), T::g_slot(x, y - 1, z), T::g_slot(x, y, z - 1)};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double res = 0.0;
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const double* Gc = s_g + gidx(r, jj) * T::NG;
        const double d0 = 0.5 * (Gc[sp[0]] - Gc[sm[0]]);
        const double d1 = 0.5 * (Gc[sp[1]] - Gc[sm[1]]);
        const double d2 = 0.5 * (Gc[sp[2]] - Gc[sm[2]]);
        res = res + (d0 * kk[0 * 3 + jj] + d1 * kk[1 * 3 + jj] + d2 * kk[2 * 3 + jj]);
      }
      vis[r] = res;
    }
  }
  __syncthreads();   // G space is reused by the face buffer (and the primitives)
  if constexpr (RS) {
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int h = t + it * T::NT;
      if (h < T::NPV) {
        double pp[5];
        cons_to_prim_fast(qq[it], g, pp);
#pragma unroll
        for (int v = 0; v < 5; ++v) s_pv[v * T::NPV + h] = pp[v];
      }
    }
    __syncthreads();
  }

  // ---- 3. inviscid faces, axis by axis ---------------------------------------
  double kl0[3], kl1[3], kl2[3];
  load_kl_own<0>(b, c, gi, own, kl0);
  load_kl_own<1>(b, c, gj, own, kl1);
  load_kl_own<2>(b, c, gk, own, kl2);
  double inv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double rvol = own ? rcp(b.vol[c] + 1e-300) : 0.0;
  face_axis<0, T, TI, TJ, TK, T::NT, LIM>(b, i0, j0, k0, s_pv, s_f, kk, kl0, own, t, x, y, z, rvol, g, inv);
  face_axis<1, T, TI, TJ, TK, T::NT, LIM>(b, i0, j0, k0, s_pv, s_f, kk, kl1, own, t, x, y, z, rvol, g, inv);
  face_axis<2, T, TI, TJ, TK, T::NT, LIM>(b, i0, j0, k0, s_pv, s_f, kk, kl2, own, t, x, y, z, rvol, g, inv);
  if (!own) return;

  // ---- 4. rhs and the RK3 combination ----------------------------------------
  double rhs[5];
  rhs[0] = inv[0] + 0.0;
  rhs[1] = inv[1] + vis[0];
  rhs[2] = inv[2] + vis[1];
  rhs[3] = inv[3] + vis[2];
  rhs[4] = inv[4] + vis[3];
  double qs[5], qq0[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    qs[v] = qin[v * N + c];
    qq0[v] = stage == 1 ? 0.0 : q0[v * N + c];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) qout[v * N + c] = rk_combine(stage, qq0[v], qs[v], dt, rhs[v]);
}

template <int TI, int TJ, int TK, int LIM, bool RS>
__device__ __forceinline__ void stage_tile_fused(const BlockDesc& b, int i0, int j0, int k0,
                                                 const double* __restrict__ qin, const double* q0, double* qout,
                                                 const Gas& gas, double dt, int stage) {
  FusedCells<TI, TJ, TK> fc;
  fused_cells<TI, TJ, TK>(b, i0, j0, k0, (int)threadIdx.x, fc);
  double qq[FusedCells<TI, TJ, TK>::NIT][5];
  fused_stage_loads<TI, TJ, TK>(b, qin, fc, qq);
  stage_tile_fused_regs<TI, TJ, TK, LIM, RS>(b, i0, j0, k0, qin, q0, qout, gas, dt, stage, fc, qq);
}

// WPE: waves per SIMD the register allocation aims at (3 for the re-staging
// form at 256 threads: three workgroups per CU, at most 168 VGPRs).
template <int TI, int TJ, int TK, int LIM, bool RS, int WPE>
__global__ void __launch_bounds__(TI * TJ * TK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_stage_fused(BlockDesc b, int stage, Gas gas, double dt, int ntj, int ntk) {
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int tk = t % ntk, tj = (t / ntk) % ntj, ti = t / (ntk * ntj);
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
  stage_tile_fused<TI, TJ, TK, LIM, RS>(b, ti * TI, tj * TJ, tk * TK, qin, b.q, qout, gas, dt, stage);
}

template <int TI, int TJ, int TK, int LIM, bool RS, int WPE>
__global__ void __launch_bounds__(TI * TJ * TK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_stage_fused_batched(const BlockDesc* __restrict__ blocks, const Tile3* __restrict__ tiles, int stage, Gas gas,
                      double dt) {
  const Tile3 tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  const BlockDesc& b = blocks[tl.block];
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
  stage_tile_fused<TI, TJ, TK, LIM, RS>(b, tl.i0, tl.j0, tl.k0, qin, b.q, qout, gas, dt, stage);
}

// ===========================================================================
// Split stage: two launches per RK3 stage (tile tables as the fused kernel).
//
//  k_visc_tensor  u, v, w, T of the tile and a one-cell cross halo staged in
//                 LDS; the viscous flux tensor G (rhs/viscous.py:30-108) of
//                 every cell, once, written to the plan's G buffer (9, N).
//  k_stage_split  primitives of the tile and a two-cell cross halo in LDS;
//                 divergence of G at the cell's six (wrapped) neighbours read
//                 from the G buffer; MUSCL-2 + Roe faces once per tile as in
//                 the fused kernel; the Shu-Osher combination.
// Every value is produced by the same expressions, on the same inputs, in
// the same order as in the fused kernel, so both kernels give the same bits;
// the split form holds no viscous tensor in LDS (43-74 KB per workgroup
// instead of 75-143 KB), so more workgroups share a CU.
// ===========================================================================
template <int TI, int TJ, int TK>
struct SplitGeom {
  static_assert(TI < WRAP_ONCE_MIN && TJ < WRAP_ONCE_MIN && TK < WRAP_ONCE_MIN, "tile longer than the single-wrap bound");
  static constexpr int NT = TI * TJ * TK;
  static constexpr int A0 = TJ * TK, A1 = TI * TK, A2 = TI * TJ;
  static constexpr int NP = NT + 4 * (A0 + A1 + A2);    // tile + two halo layers per side, per axis
  static constexpr int F0 = (TI + 1) * TJ * TK, F1 = TI * (TJ + 1) * TK, F2 = TI * TJ * (TK + 1);
  static constexpr int NF = F0 > F1 ? (F0 > F2 ? F0 : F2) : (F1 > F2 ? F1 : F2);
  static constexpr int NLDS = 5 * NP + 5 * NF;
  static constexpr int NPVS = NP;    // primitive slots (stride of a variable)

  // slot of the cell (lx, ly, lz) whose coordinate along D lies in
  // [-2, T_D + 1] and whose other two coordinates lie in the tile
  template <int D>
  __device__ static __forceinline__ int axis_slot(int lx, int ly, int lz) {
    if constexpr (D == 0) {
      const bool out = (unsigned)lx >= (unsigned)TI;
      return out ? NT + ((lx < 0 ? lx + 2 : lx - TI + 2) * TJ + ly) * TK + lz : (lx * TJ + ly) * TK + lz;
    } else if constexpr (D == 1) {
      const bool out = (unsigned)ly >= (unsigned)TJ;
      return out ? NT + 4 * A0 + ((ly < 0 ? ly + 2 : ly - TJ + 2) * TI + lx) * TK + lz : (lx * TJ + ly) * TK + lz;
    } else {
      const bool out = (unsigned)lz >= (unsigned)TK;
      return out ? NT + 4 * (A0 + A1) + ((lz < 0 ? lz + 2 : lz - TK + 2) * TI + lx) * TJ + ly
                 : (lx * TJ + ly) * TK + lz;
    }
  }
  // threads that give every axis's faces one pass (a multiple of 64)
  static constexpr int NTH1 = (NF + 63) / 64 * 64;

  // slot of a cell that lies in the tile or on the halo layers of one axis
  // (written with selects only: no divergent branches in the face loops)
  __device__ static __forceinline__ int p_slot(int x, int y, int z) {
    const bool ox = (unsigned)x >= (unsigned)TI, oy = (unsigned)y >= (unsigned)TJ, oz = (unsigned)z >= (unsigned)TK;
    const int lx = x < 0 ? x + 2 : x - TI + 2, ly = y < 0 ? y + 2 : y - TJ + 2, lz = z < 0 ? z + 2 : z - TK + 2;
    const int s_in = (x * TJ + y) * TK + z;
    const int s_x = NT + (lx * TJ + y) * TK + z;
    const int s_y = NT + 4 * A0 + (ly * TI + x) * TK + z;
    const int s_z = NT + 4 * (A0 + A1) + (lz * TI + x) * TJ + y;
    return ox ? s_x : (oy ? s_y : (oz ? s_z : s_in));
  }
  // inverse of p_slot
  __device__ static __forceinline__ int3 p_cell(int h) {
    if (h < NT) return make_int3(h / (TJ * TK), (h / TK) % TJ, h % TK);
    h -= NT;
    if (h < 4 * A0) {
      const int l = h / A0, r = h % A0;
      return make_int3(l < 2 ? l - 2 : TI + l - 2, r / TK, r % TK);
    }
    h -= 4 * A0;
    if (h < 4 * A1) {
      const int l = h / A1, r = h % A1;
      return make_int3(r / TK, l < 2 ? l - 2 : TJ + l - 2, r % TK);
    }
    h -= 4 * A1;
    const int l = h / A2, r = h % A2;
    return make_int3(r / TJ, r % TJ, l < 2 ? l - 2 : TK + l - 2);
  }
};

template <int TI, int TJ, int TK>
__device__ __forceinline__ void visc_tile(const BlockDesc& b, int i0, int j0, int k0, const double* __restrict__ qin,
                                          const Gas& gas) {
  using T = FusedGeom<TI, TJ, TK>;       // g_slot / g_cell: tile + one-cell cross halo
  __shared__ double s_w[4 * T::NG];      // u, v, w, T
  const int t = threadIdx.x;
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const int64_t si = (int64_t)nj * nk, sj = nk;
  Gas gb = gas;
  gb.mu0 = b.mu0;
  {
    constexpr int NIT = (T::NG + T::NT - 1) / T::NT;
    int cc[NIT];
    auto cells = [&](auto mod) {
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int h = t + it * T::NT;
        const Loc3 lc = T::g_cell(h < T::NG ? h : T::NG - 1);
        cc[it] = tile_cell<decltype(mod)::value>(i0 + lc.x, j0 + lc.y, k0 + lc.z, ni, nj, nk);
      }
    };
    if (b.wrap_mod) cells(std::true_type{});
    else cells(std::false_type{});
    double qq[NIT][5];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int h = t + it * T::NT;
      if (h < T::NG) {
#pragma unroll
        for (int v = 0; v < 5; ++v) qq[it][v] = qin[v * N + cc[it]];
      }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int h = t + it * T::NT;
      if (h < T::NG) {
        double pp[5];
        cons_to_prim_fast(qq[it], gas.gamma, pp);
        s_w[h] = pp[1];
        s_w[T::NG + h] = pp[2];
        s_w[2 * T::NG + h] = pp[3];
        s_w[3 * T::NG + h] = temperature_fast(pp, gb.R);
      }
    }
  }
  __syncthreads();
  const int x = t / (TJ * TK), y = (t / TK) % TJ, z = t % TK;
  const int gi = i0 + x, gj = j0 + y, gk = k0 + z;
  if (gi >= ni || gj >= nj || gk >= nk) return;
  const int64_t c = (int64_t)gi * si + (int64_t)gj * sj + gk;
  const int s0 = T::g_slot(x, y, z);
  const double pv[5] = {0.0, s_w[s0], s_w[T::NG + s0], s_w[2 * T::NG + s0], 0.0};
  const double Tc = s_w[3 * T::NG + s0];
  double d[3][4];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int sp = T::g_slot(x + (l == 0), y + (l == 1), z + (l == 2));
    const int sm = T::g_slot(x - (l == 0), y - (l == 1), z - (l == 2));
#pragma unroll
    for (int v = 0; v < 4; ++v) d[l][v] = 0.5 * (s_w[v * T::NG + sp] - s_w[v * T::NG + sm]);
  }
  double kc[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kc[m] = b.kxyz[m * N + c];
  double g9[9];
  viscous_flux_fast(pv, Tc, d, kc, gb, g9);
#pragma unroll
  for (int m = 0; m < 9; ++m) b.G[m * N + c] = g9[m];
}

template <int TI, int TJ, int TK, int NTH, int LIM>
__device__ __forceinline__ void stage_tile_split(const BlockDesc& b, int i0, int j0, int k0,
                                                 const double* __restrict__ qin, const double* q0, double* qout,
                                                 const Gas& gas, double dt, int stage) {
  using T = SplitGeom<TI, TJ, TK>;
  static_assert(NTH >= T::NT && NTH % 64 == 0, "stage kernel needs a thread per cell");
  __shared__ double lds[T::NLDS];
  double* s_pv = lds;                    // [5][NP]
  double* s_f = lds + 5 * T::NP;         // [5][NF]
  const int t = threadIdx.x;
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  const int64_t N = b.N;
  const int64_t si = (int64_t)nj * nk, sj = nk;
  const double g = gas.gamma;
  dt = dt * b.dt_scale;
  // threads t < NT own cell (x, y, z) of the tile; all NTH threads compute faces
  const int x = t / (TJ * TK), y = (t / TK) % TJ, z = t % TK;
  const int gi = i0 + x, gj = j0 + y, gk = k0 + z;
  const bool own = t < T::NT && gi < ni && gj < nj && gk < nk;
  // the thread's cell (wrapped into the block when outside it: a valid
  // address) and its staging cells
  constexpr int NIT = (T::NP + NTH - 1) / NTH;
  int c, cc[NIT];
  auto cells = [&](auto mod) {
    constexpr bool M = decltype(mod)::value;
    c = tile_cell<M>(gi, gj, gk, ni, nj, nk);
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int h = t + it * NTH;
      const int3 lc = T::p_cell(h < T::NP ? h : T::NP - 1);
      cc[it] = tile_cell<M>(i0 + lc.x, j0 + lc.y, k0 + lc.z, ni, nj, nk);
    }
  };
  if (b.wrap_mod) cells(std::true_type{});
  else cells(std::false_type{});

  // ---- 1. primitives of the tile and the two-cell cross halo ------------------
  {
    double qq[NIT][5];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int h = t + it * NTH;
      if (h < T::NP) {
#pragma unroll
        for (int v = 0; v < 5; ++v) qq[it][v] = qin[v * N + cc[it]];
      }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int h = t + it * NTH;
      if (h < T::NP) {
        double pp[5];
        cons_to_prim_fast(qq[it], g, pp);
#pragma unroll
        for (int v = 0; v < 5; ++v) s_pv[v * T::NP + h] = pp[v];
      }
    }
  }

  // ---- 2. divergence of the viscous tensor (G of the six wrapped neighbours) ---
  double kk[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) kk[m] = own ? b.kxyz[m * N + c] : 0.0;
  double kl0[3], kl1[3], kl2[3];
  load_kl_own<0>(b, c, gi, own, kl0);
  load_kl_own<1>(b, c, gj, own, kl1);
  load_kl_own<2>(b, c, gk, own, kl2);
  double vis[4] = {0.0, 0.0, 0.0, 0.0};
  if (own) {
    const int64_t cp[3] = {c + (int64_t)(wrap_p(gi, ni) - gi) * si, c + (int64_t)(wrap_p(gj, nj) - gj) * sj,
                           c + (wrap_p(gk, nk) - gk)};
    const int64_t cm[3] = {c + (int64_t)(wrap_m(gi, ni) - gi) * si, c + (int64_t)(wrap_m(gj, nj) - gj) * sj,
                           c + (wrap_m(gk, nk) - gk)};
    const double* __restrict__ Gb = b.G;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double res = 0.0;
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const double* Gc = Gb + gidx(r, jj) * N;
        const double d0 = 0.5 * (Gc[cp[0]] - Gc[cm[0]]);
        const double d1 = 0.5 * (Gc[cp[1]] - Gc[cm[1]]);
        const double d2 = 0.5 * (Gc[cp[2]] - Gc[cm[2]]);
        res = res + (d0 * kk[0 * 3 + jj] + d1 * kk[1 * 3 + jj] + d2 * kk[2 * 3 + jj]);
      }
      vis[r] = res;
    }
  }
  __syncthreads();

  // ---- 3. inviscid faces, axis by axis ---------------------------------------
  double inv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double rvol = own ? rcp(b.vol[c] + 1e-300) : 0.0;
  face_axis<0, T, TI, TJ, TK, NTH, LIM>(b, i0, j0, k0, s_pv, s_f, kk, kl0, own, t, x, y, z, rvol, g, inv);
  face_axis<1, T, TI, TJ, TK, NTH, LIM>(b, i0, j0, k0, s_pv, s_f, kk, kl1, own, t, x, y, z, rvol, g, inv);
  face_axis<2, T, TI, TJ, TK, NTH, LIM>(b, i0, j0, k0, s_pv, s_f, kk, kl2, own, t, x, y, z, rvol, g, inv);
  if (!own) return;

  // ---- 4. rhs and the RK3 combination ----------------------------------------
  double rhs[5];
  rhs[0] = inv[0] + 0.0;
  rhs[1] = inv[1] + vis[0];
  rhs[2] = inv[2] + vis[1];
  rhs[3] = inv[3] + vis[2];
  rhs[4] = inv[4] + vis[3];
  double qs[5], qq0[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    qs[v] = qin[v * N + c];
    qq0[v] = stage == 1 ? 0.0 : q0[v * N + c];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) qout[v * N + c] = rk_combine(stage, qq0[v], qs[v], dt, rhs[v]);
}

template <int TI, int TJ, int TK>
__global__ void __launch_bounds__(TI * TJ * TK) k_visc_tensor(BlockDesc b, int stage, Gas gas, int ntj, int ntk) {
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int tk = t % ntk, tj = (t / ntk) % ntj, ti = t / (ntk * ntj);
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  visc_tile<TI, TJ, TK>(b, ti * TI, tj * TJ, tk * TK, qin, gas);
}

template <int TI, int TJ, int TK, int NTH, int LIM>
__global__ void __launch_bounds__(NTH) k_stage_split(BlockDesc b, int stage, Gas gas, double dt, int ntj, int ntk) {
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int tk = t % ntk, tj = (t / ntk) % ntj, ti = t / (ntk * ntj);
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
  stage_tile_split<TI, TJ, TK, NTH, LIM>(b, ti * TI, tj * TJ, tk * TK, qin, b.q, qout, gas, dt, stage);
}

template <int TI, int TJ, int TK>
__global__ void __launch_bounds__(TI * TJ * TK) k_visc_tensor_batched(const BlockDesc* __restrict__ blocks,
                                                                      const Tile3* __restrict__ tiles, int stage,
                                                                      Gas gas) {
  const Tile3 tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  const BlockDesc& b = blocks[tl.block];
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  visc_tile<TI, TJ, TK>(b, tl.i0, tl.j0, tl.k0, qin, gas);
}

template <int TI, int TJ, int TK, int NTH, int LIM>
__global__ void __launch_bounds__(NTH) k_stage_split_batched(const BlockDesc* __restrict__ blocks,
                                                             const Tile3* __restrict__ tiles, int stage, Gas gas,
                                                             double dt) {
  const Tile3 tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  const BlockDesc& b = blocks[tl.block];
  const double* qin = stage == 1 ? b.q : (stage == 2 ? b.q1 : b.q2);
  double* qout = stage == 1 ? b.q1 : (stage == 2 ? b.q2 : b.q);
  stage_tile_split<TI, TJ, TK, NTH, LIM>(b, tl.i0, tl.j0, tl.k0, qin, b.q, qout, gas, dt, stage);
}

}  // namespace famrdp

#include "famrdp_march.h"

namespace famrdp {

// Stage-kernel variants the autotuner chooses from (ids 0-4 and 10-13
// fused, 5-9 split pair, 14- i-marching).  Fused kernel: (id, TI, TJ, TK, re-staging, waves
// per SIMD for the register allocation), one thread per cell.  Split pair:
// (id, TI, TJ, TK, threads of the stage kernel): with one thread per cell an
// axis's faces take two passes, with a thread per face (NTH1) one.
#define FAMRDP_TILE_VARIANTS(X) \
  X(0, 8, 4, 8, false, 1)       \
  X(1, 4, 8, 16, false, 1)      \
  X(2, 8, 4, 16, false, 1)      \
  X(3, 4, 8, 8, false, 1)       \
  X(4, 4, 8, 8, true, 3)        \
  X(10, 8, 4, 8, true, 3)       \
  X(11, 8, 8, 4, true, 3)       \
  X(12, 4, 4, 16, true, 3)      \
  X(13, 8, 8, 8, true, 4)
#define FAMRDP_SPLIT_VARIANTS(X)  \
  X(5, 4, 4, 16, 256)             \
  X(6, 4, 8, 8, 256)              \
  X(7, 8, 8, 8, 512)              \
  X(8, 2, 8, 16, 256)             \
  X(9, 4, 8, 8, 320)
// i-marching kernel: (id, TJ, TK, segment length L, waves per SIMD); its
// tile table holds (2 x segment index + strip bit, j0, k0) per march_tiling.
// The 10-plane segments of 18 are for blocks of about 100^3 cells, which the
// longer segments leave with fewer workgroups than the GPU holds at once.
#define FAMRDP_MARCH_VARIANTS(X) \
  X(14, 16, 16, 32, 2)           \
  X(15, 8, 32, 20, 2)            \
  X(16, 8, 32, 32, 2)            \
  X(17, 8, 32, 52, 2)            \
  X(18, 8, 32, 10, 2)
constexpr int NUM_TILE_VARIANTS = 19;

inline bool is_march_variant(int v) {
  switch (v) {
#define FAMRDP_IM(id, b, c, l, w) case id: return true;
    FAMRDP_MARCH_VARIANTS(FAMRDP_IM)
#undef FAMRDP_IM
    default: return false;
  }
}

inline void tile_dims(int v, int& ti, int& tj, int& tk) {
  switch (v) {
#define FAMRDP_TD5(id, a, b, c, n) case id: ti = a; tj = b; tk = c; return;
#define FAMRDP_TD6(id, a, b, c, rs, w) case id: ti = a; tj = b; tk = c; return;
#define FAMRDP_TDM(id, b, c, l, w) case id: ti = l; tj = b; tk = c; return;
    FAMRDP_TILE_VARIANTS(FAMRDP_TD6)
    FAMRDP_SPLIT_VARIANTS(FAMRDP_TD5)
    FAMRDP_MARCH_VARIANTS(FAMRDP_TDM)
#undef FAMRDP_TD5
#undef FAMRDP_TD6
#undef FAMRDP_TDM
    default: ti = 4; tj = 4; tk = 16;
  }
}

// One stage of one block (plain launch) with variant v; the limiter is a
// template parameter of the stage kernels (LIM_MINMOD / LIM_VANLEER).
template <int LIM>
inline void launch_stage_lim(int v, const BlockDesc& d, int stage, const Gas& gas, double dt, int ncu,
                             hipStream_t st) {
  int TI, TJ, TK;
  tile_dims(v, TI, TJ, TK);
  const int nti = (d.ni + TI - 1) / TI, ntj = (d.nj + TJ - 1) / TJ, ntk = (d.nk + TK - 1) / TK;
  const int nt = nti * ntj * ntk;
  switch (v) {
#define FAMRDP_TL(id, a, b, c, rs, w) \
  case id: k_stage_fused<a, b, c, LIM, rs, w><<<nt, a * b * c, 0, st>>>(d, stage, gas, dt, ntj, ntk); return;
    FAMRDP_TILE_VARIANTS(FAMRDP_TL)
#undef FAMRDP_TL
#define FAMRDP_SL(id, a, b, c, n)                                                       \
  case id:                                                                              \
    k_visc_tensor<a, b, c><<<nt, a * b * c, 0, st>>>(d, stage, gas, ntj, ntk);          \
    k_stage_split<a, b, c, n, LIM><<<nt, n, 0, st>>>(d, stage, gas, dt, ntj, ntk);      \
    return;
    FAMRDP_SPLIT_VARIANTS(FAMRDP_SL)
#undef FAMRDP_SL
#define FAMRDP_ML(id, b, c, l, w)                                                                \
  case id:                                                                                       \
    k_stage_march<b, c, l, LIM, w><<<nti * march_tiling(b, c, d.nj, d.nk).count(), b * c, 0, st>>>( \
        d, stage, gas, dt);                                                                      \
    return;
    FAMRDP_MARCH_VARIANTS(FAMRDP_ML)
#undef FAMRDP_ML
  }
}

inline void launch_fused(int v, const BlockDesc& d, int stage, const Gas& gas, int lim, double dt, int ncu,
                         hipStream_t st) {
  if (lim == LIM_VANLEER) launch_stage_lim<LIM_VANLEER>(v, d, stage, gas, dt, ncu, st);
  else launch_stage_lim<LIM_MINMOD>(v, d, stage, gas, dt, ncu, st);
}

// One stage of every block in one launch (two for the split pair) over a tile
// table built for v.
template <int LIM>
inline void launch_stage_batched_lim(int v, const BlockDesc* blocks, const Tile3* tiles, int ntiles, int stage,
                                     const Gas& gas, double dt, int ncu, hipStream_t st) {
  switch (v) {
#define FAMRDP_TB(id, a, b, c, rs, w)                                                                       \
  case id:                                                                                                   \
    k_stage_fused_batched<a, b, c, LIM, rs, w><<<ntiles, a * b * c, 0, st>>>(blocks, tiles, stage, gas, dt); \
    return;
    FAMRDP_TILE_VARIANTS(FAMRDP_TB)
#undef FAMRDP_TB
#define FAMRDP_SB(id, a, b, c, n)                                                                  \
  case id:                                                                                         \
    k_visc_tensor_batched<a, b, c><<<ntiles, a * b * c, 0, st>>>(blocks, tiles, stage, gas);       \
    k_stage_split_batched<a, b, c, n, LIM><<<ntiles, n, 0, st>>>(blocks, tiles, stage, gas, dt);   \
    return;
    FAMRDP_SPLIT_VARIANTS(FAMRDP_SB)
#undef FAMRDP_SB
#define FAMRDP_MB(id, b, c, l, w)                                                                      \
  case id:                                                                                              \
    k_stage_march_batched<b, c, l, LIM, w><<<ntiles, b * c, 0, st>>>(blocks, tiles, stage, gas, dt);    \
    return;
    FAMRDP_MARCH_VARIANTS(FAMRDP_MB)
#undef FAMRDP_MB
  }
}

inline void launch_fused_batched(int v, const BlockDesc* blocks, const Tile3* tiles, int ntiles, int stage,
                                 const Gas& gas, int lim, double dt, int ncu, hipStream_t st) {
  if (lim == LIM_VANLEER) launch_stage_batched_lim<LIM_VANLEER>(v, blocks, tiles, ntiles, stage, gas, dt, ncu, st);
  else launch_stage_batched_lim<LIM_MINMOD>(v, blocks, tiles, ntiles, stage, gas, dt, ncu, st);
}

}  // namespace famrdp
