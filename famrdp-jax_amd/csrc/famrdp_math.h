// Pointwise FP64 numerics of famrdp_jax's RHS path, shared by every kernel.
//
// Each function restates one reference formula in the reference's evaluation
// order (so results track the CPU restatement to rounding), written as plain
// scalar FP64 code for one cell or one face.  Citations are to the reference
// repository (famrdp_jax/...).
#pragma once

#include <hip/hip_runtime.h>

#define FD __host__ __device__ __forceinline__

namespace famrdp {

enum Limiter : int { LIM_MINMOD = 0, LIM_VANLEER = 1 };

// NaN semantics.  jnp.maximum / jnp.minimum / jnp.sign return NaN when an
// operand is NaN, while C's fmax / fmin return the other operand; the floors
// and the limiter below use these selects so a NaN state goes NaN the way
// the reference's does.
//   max_floor(x, c) = jnp.maximum(x, c) for a constant floor c (NaN x -> NaN)
FD double max_floor(double x, double c) { return x < c ? c : x; }
//   jnp.minimum(a, b) / jnp.maximum(a, b)
FD double min_nan(double a, double b) { return a != a ? a : (b != b ? b : fmin(a, b)); }
FD double max_nan(double a, double b) { return a != a ? a : (b != b ? b : fmax(a, b)); }

// physics/eos.py:12-17
FD void cons_to_prim(const double q[5], double g, double pv[5]) {
  const double rho = q[0];
  const double u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
  const double ke = 0.5 * (u * u + v * v + w * w);
  pv[0] = rho; pv[1] = u; pv[2] = v; pv[3] = w;
  pv[4] = (g - 1.0) * (q[4] - rho * ke);
}

// physics/eos.py:5-9
FD void prim_to_cons(const double pv[5], double g, double U[5]) {
  const double rho = pv[0], u = pv[1], v = pv[2], w = pv[3], p = pv[4];
  const double ke = 0.5 * (u * u + v * v + w * w);
  U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = rho * w;
  U[4] = p / (g - 1.0) + rho * ke;
}

// physics/eos.py:24-25; (T/T0)**1.5 evaluated as x*sqrt(x) (NaN for T < 0,
// as the reference's power)
FD double sutherland(double T, double mu0, double T0, double Ts) {
  const double x = T / T0;
  return mu0 * (x * sqrt(x)) * (T0 + Ts) / (T + Ts);
}

// jnp.sign: -1, 0 (signed), +1, NaN for NaN
FD double sgn(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : a); }

// physics/reconstruct/muscl2.py:5-7: s = sign(a)*max(sign(a)sign(b),0); s*min(|a|,|b|)
// (NaN in either argument gives NaN)
FD double minmod(double a, double b) {
  const double sa = sgn(a);
  const double s = sa * max_nan(sa * sgn(b), 0.0);
  return s * min_nan(fabs(a), fabs(b));
}

// physics/reconstruct/muscl2.py:10-12
FD double vanleer(double a, double b) {
  const double ab = a + b;
  const double safe = fabs(ab) > 1e-30 ? ab : 1.0;
  return (a * b > 0.0) ? (2.0 * a * b / safe) : 0.0;
}

FD double limit(int lim, double a, double b) {
  return lim == LIM_VANLEER ? vanleer(a, b) : minmod(a, b);
}

// minmod without the sign products: same-sign arguments give the smaller
// magnitude with a's sign, anything else gives zero (the reference's
// s * min(|a|,|b|) with s in {-1, 0, 1}; a zero's sign never reaches q).
// As the median of (a, b, 0), max(min(a,b), min(max(a,b),0)): both positive
// gives max(min(a,b), 0) = min(a,b), both negative max(min(a,b), max(a,b)) =
// max(a,b) (the one nearer zero), mixed signs or a zero give 0 -- the
// reference's value exactly, in four min/max, for non-NaN arguments.  A NaN
// argument is dropped here (hardware min/max return the other operand; two
// NaNs give 0); the face routine that calls it (face_flux_prim) restores the
// reference's NaN propagation for the whole four-cell stencil instead, at two
// instructions per face rather than two per limiter call.
FD double minmod_fast(double a, double b) { return fmax(fmin(a, b), fmin(fmax(a, b), 0.0)); }

// flux/roe.py:5-18
FD void physical_flux(const double U[5], const double n[3], double g, double F[5]) {
  const double rho = max_floor(U[0], 1e-300);
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

// flux/roe.py:21-85: Roe flux for a unit normal n, Harten fix on lambda1/5.
FD void roe_flux(const double UL[5], const double UR[5], const double n[3], double g,
                 double eps_entropy, double F[5]) {
  const double rhoL = max_floor(UL[0], 1e-300), rhoR = max_floor(UR[0], 1e-300);
  const double uL = UL[1] / rhoL, vL = UL[2] / rhoL, wL = UL[3] / rhoL;
  const double uR = UR[1] / rhoR, vR = UR[2] / rhoR, wR = UR[3] / rhoR;
  const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
  const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
  const double pL = (g - 1.0) * (UL[4] - rhoL * keL);
  const double pR = (g - 1.0) * (UR[4] - rhoR * keR);
  const double HL = (UL[4] + pL) / rhoL, HR = (UR[4] + pR) / rhoR;

  const double sqL = sqrt(rhoL), sqR = sqrt(rhoR);
  const double ss = max_floor(sqL + sqR, 1e-150);
  const double uh = (sqL * uL + sqR * uR) / ss;
  const double vh = (sqL * vL + sqR * vR) / ss;
  const double wh = (sqL * wL + sqR * wR) / ss;
  const double Hh = (sqL * HL + sqR * HR) / ss;
  const double q2 = uh * uh + vh * vh + wh * wh;
  const double c2 = max_floor((g - 1.0) * (Hh - 0.5 * q2), 1e-30);
  const double ch = sqrt(c2);
  const double unh = uh * n[0] + vh * n[1] + wh * n[2];
  const double rhoh = sqL * sqR;

  const double eps = eps_entropy * (fabs(unh) + ch);
  const double lam1 = unh - ch, lam5 = unh + ch;
  const double al1 = fabs(lam1) >= eps ? fabs(lam1) : 0.5 * (lam1 * lam1 / eps + eps);
  const double al5 = fabs(lam5) >= eps ? fabs(lam5) : 0.5 * (lam5 * lam5 / eps + eps);
  const double al2 = fabs(unh);

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
  const double dvt_ = dU[2] / rhoh - vh * dU[0] / rhoh;
  const double dwt = dU[3] / rhoh - wh * dU[0] / rhoh;
  const double dunf = dut * n[0] + dvt_ * n[1] + dwt * n[2];
  const double tx = dut - dunf * n[0], ty = dvt_ - dunf * n[1], tz = dwt - dunf * n[2];
  const double dvt[5] = {0.0, rhoh * tx, rhoh * ty, rhoh * tz, rhoh * (uh * tx + vh * ty + wh * tz)};

  double FL[5], FR[5];
  physical_flux(UL, n, g, FL);
  physical_flux(UR, n, g, FR);
  for (int v = 0; v < 5; ++v) {
    const double absA = al1 * a1 * R1[v] + al5 * a5 * R5[v] + al2 * (a23 * R2[v] + dvt[v]);
    F[v] = 0.5 * (FL[v] + FR[v]) - 0.5 * absA;
  }
}

// One face of rhs/invscd.py:53-97 (the MUSCL-2 slopes of the two cells either
// side, the averaged face metric and the area-weighted Roe flux).  pv0..pv3
// are the primitives of cells a-1, a, a+1, a+2; k0/k1 are kxyz[d, :] at
// cells a and a+1.  Returns F * |n| for the face between a and a+1.
FD void face_flux(const double pv0[5], const double pv1[5], const double pv2[5], const double pv3[5],
                  const double k0[3], const double k1[3], double g, int lim, double Fa[5]) {
  double pL[5], pR[5];
  for (int v = 0; v < 5; ++v) {
    const double sL = limit(lim, pv1[v] - pv0[v], pv2[v] - pv1[v]);
    const double sR = limit(lim, pv2[v] - pv1[v], pv3[v] - pv2[v]);
    pL[v] = pv1[v] + 0.5 * sL;
    pR[v] = pv2[v] - 0.5 * sR;
  }
  double UL[5], UR[5];
  prim_to_cons(pL, g, UL);
  prim_to_cons(pR, g, UR);
  double nf[3] = {0.5 * (k0[0] + k1[0]), 0.5 * (k0[1] + k1[1]), 0.5 * (k0[2] + k1[2])};
  const double nm = sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]) + 1e-300;
  const double nu[3] = {nf[0] / nm, nf[1] / nm, nf[2] / nm};
  double F[5];
  roe_flux(UL, UR, nu, g, 0.1, F);
  for (int v = 0; v < 5; ++v) Fa[v] = F[v] * nm;
}

// ---------------------------------------------------------------------------
// Stage-kernel forms.  Same mathematics as above, arranged for throughput:
// each divisor is inverted once and multiplied (x * (1/y) rounds at most one
// ulp away from x / y), and the Roe flux takes the reconstructed primitives
// directly instead of converting them to conservative variables and back
// (the reference's prim_to_cons -> roe_flux round trip is the identity up to
// rounding).
// ---------------------------------------------------------------------------

#define DF __device__ __forceinline__

// 1/x from the hardware reciprocal seed refined by two Newton-Raphson steps
// (each squares the relative error), for finite non-zero x.
DF double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// sqrt(x) from the hardware reciprocal square-root seed: one coupled
// Newton step on (sqrt, 1/(2 sqrt)) and two residual corrections.  IEEE
// special cases: +-0 and +inf are returned as they are (the seed gives inf
// and 0 there), x < 0 and NaN give NaN (the seed is NaN, so g is).
DF double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  g = fma(d, h, g);
  // one class test: -0, +0, +inf
  return __builtin_amdgcn_class(x, (1 << 5) | (1 << 6) | (1 << 9)) ? x : g;
}

// 1/sqrt(x) for x > 0 from the hardware seed: two coupled Newton steps on
// (sqrt x, 1/(2 sqrt x)); within 2 ulp of 1/sqrt(x) (checked on the GPU).
DF double rsq_acc(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  h = fma(h, r, h);
  return 2.0 * h;
}

// One Newton step on the hardware seeds (relative error 2^-24 on gfx950, so
// about 2^-47 after the step: at most ~20 ulp, measured on the GPU), for the
// face flux, whose every quotient and root enters the flux at O(1) weight.
DF double rcp1(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
DF double rsq1(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return fma(y * fma(-(x * y), y, 1.0), 0.5, y);
}

// van Leer with the division as a reciprocal (stage-kernel form)
DF double vanleer_fast(double a, double b) {
  const double ab = a + b;
  const double safe = fabs(ab) > 1e-30 ? ab : 1.0;
  const double r = 2.0 * a * b * rcp(safe);
  return (a * b > 0.0) ? r : 0.0;
}

template <int LIM>
DF double limit_fast(double a, double b) {
  if constexpr (LIM == LIM_VANLEER) return vanleer_fast(a, b);
  else return minmod_fast(a, b);
}

DF void cons_to_prim_fast(const double q[5], double g, double pv[5]) {
  const double rho = q[0];
  const double r = rcp(rho);
  const double u = q[1] * r, v = q[2] * r, w = q[3] * r;
  const double ke = 0.5 * (u * u + v * v + w * w);
  pv[0] = rho; pv[1] = u; pv[2] = v; pv[3] = w;
  pv[4] = (g - 1.0) * (q[4] - rho * ke);
}

// flux/roe.py:21-85 on left/right primitives (rho, u, v, w, p), unit normal n.
DF void roe_flux_prim(const double pl[5], const double pr[5], const double n[3], double g, double F[5]) {
  const double gm1 = g - 1.0;
  const double rgm1 = rcp(gm1);
  const double rhoL = max_floor(pl[0], 1e-300), rhoR = max_floor(pr[0], 1e-300);
  const double uL = pl[1], vL = pl[2], wL = pl[3], pL = pl[4];
  const double uR = pr[1], vR = pr[2], wR = pr[3], pR = pr[4];
  const double EL = pL * rgm1 + rhoL * (0.5 * (uL * uL + vL * vL + wL * wL));
  const double ER = pR * rgm1 + rhoR * (0.5 * (uR * uR + vR * vR + wR * wR));
  // 1/sqrt(rho) gives sqrt(rho), 1/rho and, below, 1/sqrt(rhoL rhoR)
  const double isL = rsq1(rhoL), isR = rsq1(rhoR);
  const double HL = (EL + pL) * (isL * isL), HR = (ER + pR) * (isR * isR);

  const double sqL = rhoL * isL, sqR = rhoR * isR;
  // max(sqL + sqR, 1e-150) is sqL + sqR itself: both roots are >= 1e-150
  // after the 1e-300 floors (and NaN stays NaN), so the floor never acts
  const double rs = rcp1(sqL + sqR);
  const double uh = (sqL * uL + sqR * uR) * rs;
  const double vh = (sqL * vL + sqR * vR) * rs;
  const double wh = (sqL * wL + sqR * wR) * rs;
  const double Hh = (sqL * HL + sqR * HR) * rs;
  const double q2 = uh * uh + vh * vh + wh * wh;
  const double c2 = max_floor(gm1 * (Hh - 0.5 * q2), 1e-30);
  const double isc = rsq1(c2);
  const double rc2 = isc * isc;
  const double ch = c2 * isc;
  const double unh = uh * n[0] + vh * n[1] + wh * n[2];
  const double rhoh = sqL * sqR;

  const double eps = 0.1 * (fabs(unh) + ch);
  const double reps = rcp1(eps);
  const double lam1 = unh - ch, lam5 = unh + ch;
  const double al1 = fabs(lam1) >= eps ? fabs(lam1) : 0.5 * (lam1 * lam1 * reps + eps);
  const double al5 = fabs(lam5) >= eps ? fabs(lam5) : 0.5 * (lam5 * lam5 * reps + eps);
  const double al2 = fabs(unh);

  const double d0 = rhoR - rhoL;
  const double d1 = rhoR * uR - rhoL * uL, d2 = rhoR * vR - rhoL * vL, d3 = rhoR * wR - rhoL * wL;
  const double d4 = ER - EL;
  const double dp = gm1 * (d4 - uh * d1 - vh * d2 - wh * d3 + 0.5 * q2 * d0);
  const double dun = (d1 * n[0] + d2 * n[1] + d3 * n[2]) - unh * d0;
  const double rcd = rhoh * ch * dun;
  const double a1 = (dp - rcd) * (0.5 * rc2);
  const double a5 = (dp + rcd) * (0.5 * rc2);
  const double a23 = d0 - dp * rc2;

  const double rrh = isL * isR;
  const double dut = (d1 - uh * d0) * rrh, dvt_ = (d2 - vh * d0) * rrh, dwt = (d3 - wh * d0) * rrh;
  const double dunf = dut * n[0] + dvt_ * n[1] + dwt * n[2];
  const double tx = dut - dunf * n[0], ty = dvt_ - dunf * n[1], tz = dwt - dunf * n[2];

  const double unL = uL * n[0] + vL * n[1] + wL * n[2];
  const double unR = uR * n[0] + vR * n[1] + wR * n[2];
  const double w1 = al1 * a1, w5 = al5 * a5, w2 = al2 * a23, w2r = al2 * rhoh;
  // F = 1/2 (FL + FR) - 1/2 (w1 R1 + w5 R5 + al2 (a23 R2 + dvt)), with the
  // acoustic pair gathered: w1 (u - c n) + w5 (u + c n) = (w1 + w5) u + c (w5 - w1) n
  const double mL = rhoL * unL, mR = rhoR * unR;        // mass fluxes
  const double sw = w1 + w5, dw = ch * (w5 - w1), sa = sw + w2;
  F[0] = 0.5 * ((mL + mR) - sa);
  F[1] = 0.5 * (((mL * uL + pL * n[0]) + (mR * uR + pR * n[0])) - (uh * sa + n[0] * dw + w2r * tx));
  F[2] = 0.5 * (((mL * vL + pL * n[1]) + (mR * vR + pR * n[1])) - (vh * sa + n[1] * dw + w2r * ty));
  F[3] = 0.5 * (((mL * wL + pL * n[2]) + (mR * wR + pR * n[2])) - (wh * sa + n[2] * dw + w2r * tz));
  F[4] = 0.5 * (((EL + pL) * unL + (ER + pR) * unR) -
                (Hh * sw + unh * dw + w2 * (0.5 * q2) + w2r * (uh * tx + vh * ty + wh * tz)));
}

// MUSCL-2 face states of one cell along one axis (physics/reconstruct/
// muscl2.py:18-37) from the primitives of the cell (pv) and its neighbours
// (pvm below, pvp above): with the limited slope s = lim(pv - pvm, pvp - pv),
// the left state of the cell's high face is pv + s/2 and the right state of
// its low face pv - s/2 -- the slope face_flux_prim evaluates once for each
// of the two faces, here once.  The minmod form carries the NaN of the cell
// two away from each face in, as face_flux_prim does.
template <int LIM>
DF void muscl_cell(const double pvm[5], const double pv[5], const double pvp[5], double pLhi[5], double pRlo[5]) {
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double s = limit_fast<LIM>(pv[v] - pvm[v], pvp[v] - pv[v]);
    pLhi[v] = pv[v] + 0.5 * s;
    pRlo[v] = pv[v] - 0.5 * s;
  }
  if constexpr (LIM == LIM_MINMOD) {
    pLhi[4] = fma(0.0, pvm[4], pLhi[4]);
    pRlo[4] = fma(0.0, pvp[4], pRlo[4]);
  }
}

// The reconstructed states enter the face flux as computed values: every
// kernel form (faces reconstructed in place, or from states kept from the
// previous plane) then compiles the same flux code -- the same contractions
// of products into fused multiply-adds -- so all forms give the same bits.
DF void face_state_fence(double x[5]) {
#pragma unroll
  for (int v = 0; v < 5; ++v) asm volatile("" : "+v"(x[v]));
}

// Area-weighted Roe flux of one face from its reconstructed left/right
// primitives and the face normal nf = (kxyz[d, :] at the two cells) / 2
// (rhs/invscd.py:53-97).
DF void face_flux_n(const double pL[5], const double pR[5], double nf0, double nf1, double nf2, double g,
                    double Fa[5]) {
  // |n| + 1e-300 and its reciprocal from one 1/sqrt (|n| = 0: 1e-300, 1e300)
  const double nn = nf0 * nf0 + nf1 * nf1 + nf2 * nf2;
  const double isn = rsq1(nn);
  const bool pos = nn > 0.0;
  const double nm = (pos ? nn * isn : 0.0) + 1e-300;
  const double rn = pos ? isn : 1e300;
  const double nu[3] = {nf0 * rn, nf1 * rn, nf2 * rn};
  double F[5];
  roe_flux_prim(pL, pR, nu, g, F);
#pragma unroll
  for (int v = 0; v < 5; ++v) Fa[v] = F[v] * nm;
}

// ... from kxyz[d, :] at the cells either side
DF void face_flux_lr(const double pL[5], const double pR[5], const double k0[3], const double k1[3], double g,
                     double Fa[5]) {
  face_flux_n(pL, pR, 0.5 * (k0[0] + k1[0]), 0.5 * (k0[1] + k1[1]), 0.5 * (k0[2] + k1[2]), g, Fa);
}

// ... from kxyz[d, :] / 2 at the cells either side: k0h + k1h is
// 0.5 (k0 + k1) exactly (halving is exact and commutes with rounding), so
// the same bits as face_flux_lr
DF void face_flux_lr_h(const double pL[5], const double pR[5], const double k0h[3], const double k1h[3], double g,
                       double Fa[5]) {
  face_flux_n(pL, pR, k0h[0] + k1h[0], k0h[1] + k1h[1], k0h[2] + k1h[2], g, Fa);
}


// The reconstructed states of one face from the primitives of cells a-1, a,
// a+1, a+2 (rhs/invscd.py:53-97), on the stage-kernel forms.
template <int LIM>
DF void face_states_prim(const double pv0[5], const double pv1[5], const double pv2[5], const double pv3[5],
                         double pL[5], double pR[5]) {
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double sL = limit_fast<LIM>(pv1[v] - pv0[v], pv2[v] - pv1[v]);
    const double sR = limit_fast<LIM>(pv2[v] - pv1[v], pv3[v] - pv2[v]);
    pL[v] = pv1[v] + 0.5 * sL;
    pR[v] = pv2[v] - 0.5 * sR;
  }
  if constexpr (LIM == LIM_MINMOD) {
    // The reference's minmod is NaN when a slope is, which makes the face
    // flux NaN whenever any primitive of the four cells is NaN (a NaN in
    // pL or pR reaches every flux component through the pressure jump).
    // minmod_fast drops NaN, so the NaN of the outer cells is carried in
    // here: a cell with any NaN primitive has a NaN pressure (p is computed
    // from all five conservative variables), and 0 * p is +-0 otherwise.
    // (van Leer selects 0 for a NaN slope in the reference as well.)
    pL[4] = fma(0.0, pv0[4], pL[4]);
    pR[4] = fma(0.0, pv3[4], pR[4]);
  }
  face_state_fence(pL);
  face_state_fence(pR);
}

// One face from the primitives of cells a-1 .. a+2 and kxyz[d, :] at cells a
// and a+1.
template <int LIM>
DF void face_flux_prim(const double pv0[5], const double pv1[5], const double pv2[5], const double pv3[5],
                       const double k0[3], const double k1[3], double g, double Fa[5]) {
  double pL[5], pR[5];
  face_states_prim<LIM>(pv0, pv1, pv2, pv3, pL, pR);
  face_flux_lr(pL, pR, k0, k1, g, Fa);
}

// ... from kxyz[d, :] / 2 at cells a and a+1 (face_flux_lr_h: the same bits)
template <int LIM>
DF void face_flux_prim_h(const double pv0[5], const double pv1[5], const double pv2[5], const double pv3[5],
                         const double k0h[3], const double k1h[3], double g, double Fa[5]) {
  double pL[5], pR[5];
  face_states_prim<LIM>(pv0, pv1, pv2, pv3, pL, pR);
  face_flux_lr_h(pL, pR, k0h, k1h, g, Fa);
}

struct Gas {
  double gamma, R, mu0, T0, Ts, Pr_lam;
};

// rhs/viscous.py:30-108 at one cell: given the primitives of the cell, its
// central differences d[l][s] = 0.5*(s(+e_l) - s(-e_l)) of s in (u,v,w,T) and
// kxyz at the cell, produce the viscous flux tensor G = (tau00, tau01, tau02,
// tau11, tau12, tau22, E0, E1, E2) with E_j = u_i tau_ij - heat_j.
FD void viscous_flux(const double pv[5], double T, const double d[3][4], const double k[9],
                     const Gas& gas, double G[9]) {
  const double mu = sutherland(T, gas.mu0, gas.T0, gas.Ts);
  const double cp = gas.gamma * gas.R / (gas.gamma - 1.0);
  const double kappa = cp * mu / gas.Pr_lam;
  double gr[4][3];  // gr[s][m] = d s / d x_m = sum_l d[l][s] * k[l][m]
  for (int s = 0; s < 4; ++s)
    for (int m = 0; m < 3; ++m)
      gr[s][m] = d[0][s] * k[0 * 3 + m] + d[1][s] * k[1 * 3 + m] + d[2][s] * k[2 * 3 + m];
  const double divu = gr[0][0] + gr[1][1] + gr[2][2];
  const double t00 = 2 * mu * (gr[0][0] - divu / 3.0);
  const double t01 = mu * (gr[0][1] + gr[1][0]);
  const double t02 = mu * (gr[0][2] + gr[2][0]);
  const double t11 = 2 * mu * (gr[1][1] - divu / 3.0);
  const double t12 = mu * (gr[1][2] + gr[2][1]);
  const double t22 = 2 * mu * (gr[2][2] - divu / 3.0);
  const double u = pv[1], v = pv[2], w = pv[3];
  G[0] = t00; G[1] = t01; G[2] = t02; G[3] = t11; G[4] = t12; G[5] = t22;
  G[6] = u * t00 + v * t01 + w * t02 - (-kappa * gr[3][0]);
  G[7] = u * t01 + v * t11 + w * t12 - (-kappa * gr[3][1]);
  G[8] = u * t02 + v * t12 + w * t22 - (-kappa * gr[3][2]);
}

// viscous_flux on the stage-kernel forms (reciprocals, x*sqrt(x) for x^1.5).
DF void viscous_flux_fast(const double pv[5], double T, const double d[3][4], const double k[9], const Gas& gas,
                          double G[9]) {
  const double x = T * rcp(gas.T0);
  const double mu = gas.mu0 * (x * fsqrt(x)) * (gas.T0 + gas.Ts) * rcp(T + gas.Ts);
  const double kappa = gas.gamma * gas.R * rcp(gas.gamma - 1.0) * mu * rcp(gas.Pr_lam);
  double gr[4][3];
  for (int s = 0; s < 4; ++s)
    for (int m = 0; m < 3; ++m)
      gr[s][m] = d[0][s] * k[0 * 3 + m] + d[1][s] * k[1 * 3 + m] + d[2][s] * k[2 * 3 + m];
  const double divu3 = (gr[0][0] + gr[1][1] + gr[2][2]) * (1.0 / 3.0);
  const double t00 = 2 * mu * (gr[0][0] - divu3);
  const double t01 = mu * (gr[0][1] + gr[1][0]);
  const double t02 = mu * (gr[0][2] + gr[2][0]);
  const double t11 = 2 * mu * (gr[1][1] - divu3);
  const double t12 = mu * (gr[1][2] + gr[2][1]);
  const double t22 = 2 * mu * (gr[2][2] - divu3);
  const double u = pv[1], v = pv[2], w = pv[3];
  G[0] = t00; G[1] = t01; G[2] = t02; G[3] = t11; G[4] = t12; G[5] = t22;
  G[6] = u * t00 + v * t01 + w * t02 + kappa * gr[3][0];
  G[7] = u * t01 + v * t11 + w * t12 + kappa * gr[3][1];
  G[8] = u * t02 + v * t12 + w * t22 + kappa * gr[3][2];
}

DF double temperature_fast(const double pv[5], double R) { return pv[4] * rcp(max_floor(pv[0] * R, 1e-300)); }

// T = p / max(rho R, 1e-300)  (rhs/viscous.py:36)
FD double temperature(const double pv[5], double R) { return pv[4] / max_floor(pv[0] * R, 1e-300); }

// Tensor component index of G for (row r in 0..3, column j): rows 0..2 are
// tau_rj, row 3 is E_j.
FD int gidx(int r, int j) {
  if (r == 3) return 6 + j;
  const int a = r < j ? r : j, b = r < j ? j : r;
  return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

}  // namespace famrdp
