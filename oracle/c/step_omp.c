/*
 * Oracle (C + OpenMP): famrdp_jax's TVD-RK3 step for one block whose faces
 * are CUT1TO1 connections to the block itself (periodic in the cut axes),
 * restated from the numpy oracle (oracle/numerics.py, oracle/blocks.py) op
 * for op and in the same evaluation order, for the CPU baseline on all host
 * cores.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py's CPU-baseline leg and by the
 * test that pins it against the numpy oracle; the product never uses it.
 * Build: oracle/c/Makefile (gcc -O2 -fopenmp -ffp-contract=off, so no product
 * is fused into an FMA that numpy evaluates as two roundings).
 *
 * Reference: famrdp_jax/solver/step.py:18-56 (step), mesh/halo.py:56-92
 * (halo), rhs/invscd.py:18-133, physics/reconstruct/muscl2.py:5-37,
 * flux/roe.py:5-85, rhs/viscous.py:12-119, physics/eos.py:5-25.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double gamma, R, mu0, T0, Ts, Pr;
} gas_t;

static double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

/* muscl2.py:5-7 */
static double minmod(double a, double b) {
  const double s = sgn(a) * fmax(sgn(a) * sgn(b), 0.0);
  return s * fmin(fabs(a), fabs(b));
}

/* muscl2.py:10-12 */
static double vanleer(double a, double b) {
  const double safe = fabs(a + b) > 1e-30 ? a + b : 1.0;
  return a * b > 0.0 ? 2.0 * a * b / safe : 0.0;
}

/* eos.py:5-9 */
static void prim_to_cons(const double p[5], double g, double U[5]) {
  const double ke = 0.5 * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  U[0] = p[0];
  U[1] = p[0] * p[1];
  U[2] = p[0] * p[2];
  U[3] = p[0] * p[3];
  U[4] = p[4] / (g - 1.0) + p[0] * ke;
}

/* roe.py:5-18 */
static void physical_flux(const double U[5], const double n[3], double g, double F[5]) {
  const double rho = fmax(U[0], 1e-300);
  const double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
  const double ke = 0.5 * (u * u + v * v + w * w);
  const double p = (g - 1.0) * (U[4] - rho * ke);
  const double un = u * n[0] + v * n[1] + w * n[2];
  const double H = (U[4] + p) / rho;
  F[0] = rho * un;
  F[1] = rho * u * un + p * n[0];
  F[2] = rho * v * un + p * n[1];
  F[3] = rho * w * un + p * n[2];
  F[4] = rho * H * un;
}

static double harten(double lam, double eps) {
  return fabs(lam) >= eps ? fabs(lam) : 0.5 * (lam * lam / eps + eps);
}

/* roe.py:21-85 */
static void roe_flux(const double UL[5], const double UR[5], const double n[3], double g, double F[5]) {
  const double rhoL = fmax(UL[0], 1e-300), rhoR = fmax(UR[0], 1e-300);
  const double uL = UL[1] / rhoL, vL = UL[2] / rhoL, wL = UL[3] / rhoL;
  const double uR = UR[1] / rhoR, vR = UR[2] / rhoR, wR = UR[3] / rhoR;
  const double keL = 0.5 * (uL * uL + vL * vL + wL * wL), keR = 0.5 * (uR * uR + vR * vR + wR * wR);
  const double pL = (g - 1.0) * (UL[4] - rhoL * keL), pR = (g - 1.0) * (UR[4] - rhoR * keR);
  const double HL = (UL[4] + pL) / rhoL, HR = (UR[4] + pR) / rhoR;
  const double sqL = sqrt(rhoL), sqR = sqrt(rhoR);
  const double ss = fmax(sqL + sqR, 1e-150);
  const double uh = (sqL * uL + sqR * uR) / ss;
  const double vh = (sqL * vL + sqR * vR) / ss;
  const double wh = (sqL * wL + sqR * wR) / ss;
  const double Hh = (sqL * HL + sqR * HR) / ss;
  const double q2 = uh * uh + vh * vh + wh * wh;
  const double c2 = fmax((g - 1.0) * (Hh - 0.5 * q2), 1e-30);
  const double ch = sqrt(c2);
  const double unh = uh * n[0] + vh * n[1] + wh * n[2];
  const double rhoh = sqL * sqR;
  const double eps = 0.1 * (fabs(unh) + ch);
  const double al1 = harten(unh - ch, eps), al2 = fabs(unh), al5 = harten(unh + ch, eps);
  double dU[5];
  for (int v = 0; v < 5; ++v) dU[v] = UR[v] - UL[v];
  const double dp = (g - 1.0) * (dU[4] - uh * dU[1] - vh * dU[2] - wh * dU[3] + 0.5 * q2 * dU[0]);
  const double dun = (dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2]) - unh * dU[0];
  const double a1 = (dp - rhoh * ch * dun) / (2.0 * c2);
  const double a5 = (dp + rhoh * ch * dun) / (2.0 * c2);
  const double a23 = dU[0] - dp / c2;
  const double R1[5] = {1.0, uh - ch * n[0], vh - ch * n[1], wh - ch * n[2], Hh - ch * unh};
  const double R5[5] = {1.0, uh + ch * n[0], vh + ch * n[1], wh + ch * n[2], Hh + ch * unh};
  const double R2[5] = {1.0, uh, vh, wh, 0.5 * q2};
  const double dut = dU[1] / rhoh - uh * dU[0] / rhoh;
  const double dvt = dU[2] / rhoh - vh * dU[0] / rhoh;
  const double dwt = dU[3] / rhoh - wh * dU[0] / rhoh;
  const double dunf = dut * n[0] + dvt * n[1] + dwt * n[2];
  const double tx = dut - dunf * n[0], ty = dvt - dunf * n[1], tz = dwt - dunf * n[2];
  const double D[5] = {0.0, rhoh * tx, rhoh * ty, rhoh * tz, rhoh * (uh * tx + vh * ty + wh * tz)};
  double FL[5], FR[5];
  physical_flux(UL, n, g, FL);
  physical_flux(UR, n, g, FR);
  for (int v = 0; v < 5; ++v) {
    const double absA = al1 * a1 * R1[v] + al5 * a5 * R5[v] + al2 * (a23 * R2[v] + D[v]);
    F[v] = 0.5 * (FL[v] + FR[v]) - 0.5 * absA;
  }
}

/* invscd.py:18-133: adds the inviscid RHS of every cell line along `axis`
 * (0, 1, 2) to rhs; pv holds the primitives (5, N). */
static void invscd_axis(const double* pv, const double* kxyz, const double* vol, int ni, int nj, int nk, int axis,
                        double g, int lim, double* rhs) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int n[3] = {ni, nj, nk};
  const int64_t st[3] = {(int64_t)nj * nk, nk, 1};
  const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
  const int nd = n[axis];
  if (nd < 4) return;
#pragma omp parallel
  {
    double* F = (double*)malloc(sizeof(double) * 5 * (size_t)(nd - 3));
#pragma omp for collapse(2) schedule(static)
    for (int a = 0; a < n[o1]; ++a)
      for (int b = 0; b < n[o2]; ++b) {
        const int64_t base = a * st[o1] + b * st[o2];
        for (int f = 0; f <= nd - 4; ++f) {
          const int64_t c0 = base + (int64_t)f * st[axis];
          double pL[5], pR[5];
          for (int v = 0; v < 5; ++v) {
            const double* p = pv + v * N;
            const double pmm = p[c0], pc = p[c0 + st[axis]], pp1 = p[c0 + 2 * st[axis]], pp2 = p[c0 + 3 * st[axis]];
            const double sL = lim ? vanleer(pc - pmm, pp1 - pc) : minmod(pc - pmm, pp1 - pc);
            const double sR = lim ? vanleer(pp1 - pc, pp2 - pp1) : minmod(pp1 - pc, pp2 - pp1);
            pL[v] = pc + 0.5 * sL;
            pR[v] = pp1 - 0.5 * sR;
          }
          double UL[5], UR[5];
          prim_to_cons(pL, g, UL);
          prim_to_cons(pR, g, UR);
          const int64_t cl = c0 + st[axis], cr = c0 + 2 * st[axis];
          double nf[3];
          for (int m = 0; m < 3; ++m) nf[m] = 0.5 * (kxyz[(axis * 3 + m) * N + cl] + kxyz[(axis * 3 + m) * N + cr]);
          const double nm = sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]) + 1e-300;
          const double nu[3] = {nf[0] / nm, nf[1] / nm, nf[2] / nm};
          double Ff[5];
          roe_flux(UL, UR, nu, g, Ff);
          for (int v = 0; v < 5; ++v) F[v * (nd - 3) + f] = Ff[v] * nm;
        }
        for (int c = 2; c <= nd - 3; ++c) {
          const int64_t cc = base + (int64_t)c * st[axis];
          for (int v = 0; v < 5; ++v) {
            const double div = F[v * (nd - 3) + c - 1] - F[v * (nd - 3) + c - 2];
            rhs[v * N + cc] = rhs[v * N + cc] + -div / (vol[cc] + 1e-300);
          }
        }
      }
    free(F);
  }
}

/* 0.5 * (f(+1) - f(-1)) along axis with periodic wrap of the padded array
 * (viscous.py:12-16, jnp.roll). */
static double central2(const double* f, int i, int j, int k, int ni, int nj, int nk, int axis) {
  const int64_t si = (int64_t)nj * nk, sj = nk;
  int ip = i, im = i, jp = j, jm = j, kp = k, km = k;
  if (axis == 0) { ip = i == ni - 1 ? 0 : i + 1; im = i == 0 ? ni - 1 : i - 1; }
  if (axis == 1) { jp = j == nj - 1 ? 0 : j + 1; jm = j == 0 ? nj - 1 : j - 1; }
  if (axis == 2) { kp = k == nk - 1 ? 0 : k + 1; km = k == 0 ? nk - 1 : k - 1; }
  return 0.5 * (f[ip * si + jp * sj + kp] - f[im * si + jm * sj + km]);
}

/* viscous.py:19-119, added to rhs (components 1-4; component 0 adds 0). */
static void viscous(const double* pv, const double* kxyz, int ni, int nj, int nk, const gas_t* gas, double* rhs,
                    double* work /* 17 N */) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t si = (int64_t)nj * nk, sj = nk;
  double* T = work;
  double* mu = work + N;
  double* kap = work + 2 * N;
  double* G = work + 3 * N;  /* 12 N: tau[i][j] (9) then the energy flux (3) */
  const double cp = gas->gamma * gas->R / (gas->gamma - 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    T[c] = pv[4 * N + c] / fmax(pv[c] * gas->R, 1e-300);
    mu[c] = gas->mu0 * pow(T[c] / gas->T0, 1.5) * (gas->T0 + gas->Ts) / (T[c] + gas->Ts);
    kap[c] = cp * mu[c] / gas->Pr;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j)
      for (int k = 0; k < nk; ++k) {
        const int64_t c = i * si + j * sj + k;
        const double* f4[4] = {pv + N, pv + 2 * N, pv + 3 * N, T};
        double gr[4][3];
        for (int s = 0; s < 4; ++s) {
          const double d0 = central2(f4[s], i, j, k, ni, nj, nk, 0);
          const double d1 = central2(f4[s], i, j, k, ni, nj, nk, 1);
          const double d2 = central2(f4[s], i, j, k, ni, nj, nk, 2);
          for (int m = 0; m < 3; ++m)
            gr[s][m] = d0 * kxyz[(0 * 3 + m) * N + c] + d1 * kxyz[(1 * 3 + m) * N + c] + d2 * kxyz[(2 * 3 + m) * N + c];
        }
        const double m_ = mu[c];
        const double div = gr[0][0] + gr[1][1] + gr[2][2];
        double tau[3][3];
        tau[0][0] = 2 * m_ * (gr[0][0] - div / 3.0);
        tau[0][1] = m_ * (gr[0][1] + gr[1][0]);
        tau[0][2] = m_ * (gr[0][2] + gr[2][0]);
        tau[1][0] = m_ * (gr[0][1] + gr[1][0]);
        tau[1][1] = 2 * m_ * (gr[1][1] - div / 3.0);
        tau[1][2] = m_ * (gr[1][2] + gr[2][1]);
        tau[2][0] = m_ * (gr[0][2] + gr[2][0]);
        tau[2][1] = m_ * (gr[1][2] + gr[2][1]);
        tau[2][2] = 2 * m_ * (gr[2][2] - div / 3.0);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) G[(a * 3 + b) * N + c] = tau[a][b];
        const double u = pv[N + c], v = pv[2 * N + c], w = pv[3 * N + c];
        for (int jj = 0; jj < 3; ++jj) {
          const double heat = -kap[c] * gr[3][jj];
          G[(9 + jj) * N + c] = u * tau[0][jj] + v * tau[1][jj] + w * tau[2][jj] - heat;
        }
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j)
      for (int k = 0; k < nk; ++k) {
        const int64_t c = i * si + j * sj + k;
        for (int r = 0; r < 4; ++r) {   /* momentum 0-2 (rows of tau), energy */
          double res = 0.0;
          for (int jj = 0; jj < 3; ++jj) {
            const double* f = G + (r < 3 ? r * 3 + jj : 9 + jj) * N;
            const double d0 = central2(f, i, j, k, ni, nj, nk, 0);
            const double d1 = central2(f, i, j, k, ni, nj, nk, 1);
            const double d2 = central2(f, i, j, k, ni, nj, nk, 2);
            res = res + (d0 * kxyz[(0 * 3 + jj) * N + c] + d1 * kxyz[(1 * 3 + jj) * N + c] +
                         d2 * kxyz[(2 * 3 + jj) * N + c]);
          }
          rhs[(1 + r) * N + c] = rhs[(1 + r) * N + c] + res;
        }
        rhs[c] = rhs[c] + 0.0;
      }
}

/* rhs = inviscid + viscous of state q (step.py:10-15). */
static void rhs_block(const double* q, const double* kxyz, const double* vol, int ni, int nj, int nk,
                      const gas_t* gas, int lim, double* rhs, double* pv, double* work) {
  const int64_t N = (int64_t)ni * nj * nk;
  const double g = gas->gamma;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {     /* eos.py:12-17 */
    const double rho = q[c];
    const double u = q[N + c] / rho, v = q[2 * N + c] / rho, w = q[3 * N + c] / rho;
    const double ke = 0.5 * (u * u + v * v + w * w);
    pv[c] = rho; pv[N + c] = u; pv[2 * N + c] = v; pv[3 * N + c] = w;
    pv[4 * N + c] = (g - 1.0) * (q[4 * N + c] - rho * ke);
  }
  double* inv = work + 17 * N;          /* 5 N */
  memset(inv, 0, sizeof(double) * 5 * N);
  for (int axis = 0; axis < 3; ++axis) invscd_axis(pv, kxyz, vol, ni, nj, nk, axis, g, lim, inv);
  memset(rhs, 0, sizeof(double) * 5 * N);
  viscous(pv, kxyz, ni, nj, nk, gas, rhs, work);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < 5 * N; ++c) rhs[c] = inv[c] + rhs[c];
}

/* One step in place.  cuts: ncut pairs (target face, source face), faces
 * 0..5 = IMIN, IMAX, JMIN, JMAX, KMIN, KMAX, applied in order from the
 * pre-exchange state (halo.py:56-92); ghost = 2.  gas6: gamma, R, mu0, T0,
 * Ts, Pr_lam.  limiter: 0 minmod, 1 van Leer.  nthreads <= 0: OpenMP default.
 * Returns 0, or -1 on bad arguments or allocation failure. */
int oracle_step_self_block(double* q, const double* kxyz, const double* vol, int ni, int nj, int nk,
                           const int* cuts, int ncut, const double* gas6, double dt, int limiter, int nthreads) {
  if (!q || !kxyz || !vol || !gas6 || ni < 4 || nj < 4 || nk < 4 || (ncut > 0 && !cuts)) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const gas_t gas = {gas6[0], gas6[1], gas6[2], gas6[3], gas6[4], gas6[5]};
  const int64_t N = (int64_t)ni * nj * nk;
  const int n[3] = {ni, nj, nk};
  const int64_t st[3] = {(int64_t)nj * nk, nk, 1};
  double* buf = (double*)malloc(sizeof(double) * (size_t)(5 * N * 6 + 22 * N));
  if (!buf) return -1;
  double* src = buf;                    /* pre-exchange copy */
  double* q1 = buf + 5 * N;
  double* q2 = buf + 10 * N;
  double* L = buf + 15 * N;
  double* pv = buf + 20 * N;
  double* q0 = buf + 25 * N;            /* state after the halo */
  double* work = buf + 30 * N;          /* 22 N */
  memcpy(src, q, sizeof(double) * 5 * N);
  memcpy(q0, q, sizeof(double) * 5 * N);
  for (int e = 0; e < ncut; ++e) {      /* halo.py:19-33, 76-92 */
    const int tf = cuts[2 * e], sf = cuts[2 * e + 1];
    const int ax = tf / 2;
    if (sf / 2 != ax) { free(buf); return -1; }
    const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
    const int t0 = (tf % 2 == 0) ? 0 : n[ax] - 2;
    const int s0 = (sf % 2 == 0) ? 2 : n[ax] - 4;
    for (int v = 0; v < 5; ++v)
      for (int l = 0; l < 2; ++l)
        for (int a = 0; a < n[o1]; ++a)
          for (int b = 0; b < n[o2]; ++b) {
            const int64_t off = a * st[o1] + b * st[o2];
            q0[v * N + (t0 + l) * st[ax] + off] = src[v * N + (s0 + l) * st[ax] + off];
          }
  }
  rhs_block(q0, kxyz, vol, ni, nj, nk, &gas, limiter, L, pv, work);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < 5 * N; ++c) q1[c] = q0[c] + dt * L[c];
  rhs_block(q1, kxyz, vol, ni, nj, nk, &gas, limiter, L, pv, work);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < 5 * N; ++c) q2[c] = 0.75 * q0[c] + 0.25 * (q1[c] + dt * L[c]);
  rhs_block(q2, kxyz, vol, ni, nj, nk, &gas, limiter, L, pv, work);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < 5 * N; ++c) q[c] = (1.0 / 3.0) * q0[c] + (2.0 / 3.0) * (q2[c] + dt * L[c]);
  free(buf);
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
