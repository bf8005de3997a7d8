// Hand-written reverse-mode derivatives (vector-Jacobian products) of the
// pointwise numerics in famrdp_math.h, for the adjoint of the TVD-RK3 step.
//
// Branches are taken the way jax.grad takes them for the reference code
// (famrdp_jax/...): jnp.where sends the cotangent to the selected branch,
// jnp.maximum/minimum split it 1/2-1/2 on exact ties, jnp.abs has derivative
// select(x >= 0, 1, -1), jnp.sign has derivative 0.  The metric terms (face
// normals, kxyz, vol) are constants of the differentiation.
//
// Functions are host+device so the reverse sweeps can be checked on the CPU;
// on the GPU the reciprocals and square roots use the stage-kernel forms.
#pragma once

#include "famrdp_math.h"

namespace famrdp {

FD double rcp_hd(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rcp(x);
#else
  return 1.0 / x;
#endif
}

FD double sqrt_hd(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return fsqrt(x);
#else
  return sqrt(x);
#endif
}

// d|x|/dx as JAX takes it
FD double dabs(double x) { return x >= 0.0 ? 1.0 : -1.0; }
// d max(x, c)/dx for a constant floor c (ties split 1/2)
FD double dmax_floor(double x, double c) { return x > c ? 1.0 : (x == c ? 0.5 : 0.0); }

// cons_to_prim (physics/eos.py:12-17) reverse: lp = cotangent of
// (rho, u, v, w, p), q the conservative state; accumulates into lq.
FD void cons_to_prim_vjp(const double q[5], double g, const double lp[5], double lq[5]) {
  const double rho = q[0];
  const double r = rcp_hd(rho);
  const double u = q[1] * r, v = q[2] * r, w = q[3] * r;
  const double ke = 0.5 * (u * u + v * v + w * w);
  const double gm1 = g - 1.0;
  // p = gm1 (q4 - rho ke): dp/dq4 = gm1, dp/drho = -gm1 ke, dp/du = -gm1 rho u
  const double lu = lp[1] - gm1 * rho * u * lp[4];
  const double lv = lp[2] - gm1 * rho * v * lp[4];
  const double lw = lp[3] - gm1 * rho * w * lp[4];
  lq[4] += gm1 * lp[4];
  lq[1] += lu * r;
  lq[2] += lv * r;
  lq[3] += lw * r;
  lq[0] += lp[0] - gm1 * ke * lp[4] - (u * lu + v * lv + w * lw) * r;
}

// Limiter reverse: s = lim(A, B) (physics/reconstruct/muscl2.py:5-12).
template <int LIM>
FD void limiter_vjp_t(double A, double B, double ls, double& lA, double& lB) {
  if constexpr (LIM == LIM_VANLEER) {
    if (A * B > 0.0) {
      const double ab = A + B;
      const bool use = fabs(ab) > 1e-30;
      const double den = use ? ab : 1.0;
      const double rd = rcp_hd(den);
      const double s = 2.0 * A * B * rd;
      lA = ls * (2.0 * B * rd - (use ? s * rd : 0.0));
      lB = ls * (2.0 * A * rd - (use ? s * rd : 0.0));
    } else {
      lA = 0.0; lB = 0.0;
    }
  } else {
    // minmod: s = S min(|A|, |B|), S = sign(A) max(sign(A) sign(B), 0), dS = 0.
    // S is nonzero only when A and B have one strict sign, and then
    // S |.|'(A) = S |.|'(B) = 1: the cotangent goes to the smaller magnitude
    // (split 1/2 on a tie).
    const bool same = (A > 0.0 && B > 0.0) || (A < 0.0 && B < 0.0);
    const double xa = fabs(A), xb = fabs(B);
    const double wa = xa < xb ? 1.0 : (xa == xb ? 0.5 : 0.0);
    const double wb = xb < xa ? 1.0 : (xa == xb ? 0.5 : 0.0);
    lA = same ? ls * wa : 0.0;
    lB = same ? ls * wb : 0.0;
  }
}

FD void limiter_vjp(int lim, double A, double B, double ls, double& lA, double& lB) {
  if (lim == LIM_VANLEER) limiter_vjp_t<LIM_VANLEER>(A, B, ls, lA, lB);
  else limiter_vjp_t<LIM_MINMOD>(A, B, ls, lA, lB);
}

// Forward Roe flux on primitives (identical arithmetic to roe_flux_prim) with
// the reverse sweep: given lF (cotangent of F), accumulate into lpl, lpr.
FD void roe_flux_prim_vjp(const double pl[5], const double pr[5], const double n[3], double g, const double lF[5],
                          double lpl[5], double lpr[5]) {
  // ---------------- forward (kept intermediates) ----------------
  const double gm1 = g - 1.0;
  const double rgm1 = rcp_hd(gm1);
  const double rhoL = max_floor(pl[0], 1e-300), rhoR = max_floor(pr[0], 1e-300);
  const double uL = pl[1], vL = pl[2], wL = pl[3], pL = pl[4];
  const double uR = pr[1], vR = pr[2], wR = pr[3], pR = pr[4];
  const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
  const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
  const double EL = pL * rgm1 + rhoL * keL, ER = pR * rgm1 + rhoR * keR;
  const double irL = rcp_hd(rhoL), irR = rcp_hd(rhoR);
  const double HL = (EL + pL) * irL, HR = (ER + pR) * irR;
  const double sqL = sqrt_hd(rhoL), sqR = sqrt_hd(rhoR);
  const double S = sqL + sqR;
  const double Sf = max_floor(S, 1e-150);
  const double rs = rcp_hd(Sf);
  const double nuh = sqL * uL + sqR * uR, nvh = sqL * vL + sqR * vR, nwh = sqL * wL + sqR * wR;
  const double nHh = sqL * HL + sqR * HR;
  const double uh = nuh * rs, vh = nvh * rs, wh = nwh * rs, Hh = nHh * rs;
  const double q2 = uh * uh + vh * vh + wh * wh;
  const double c2raw = gm1 * (Hh - 0.5 * q2);
  const double c2 = max_floor(c2raw, 1e-30);
  const double rc2 = rcp_hd(c2);
  const double ch = sqrt_hd(c2);
  const double unh = uh * n[0] + vh * n[1] + wh * n[2];
  const double rhoh = sqL * sqR;
  const double eps = 0.1 * (fabs(unh) + ch);
  const double reps = rcp_hd(eps);
  const double lam1 = unh - ch, lam5 = unh + ch;
  const bool br1 = fabs(lam1) >= eps, br5 = fabs(lam5) >= eps;
  const double al1 = br1 ? fabs(lam1) : 0.5 * (lam1 * lam1 * reps + eps);
  const double al5 = br5 ? fabs(lam5) : 0.5 * (lam5 * lam5 * reps + eps);
  const double al2 = fabs(unh);
  const double d0 = rhoR - rhoL;
  const double d1 = rhoR * uR - rhoL * uL, d2 = rhoR * vR - rhoL * vL, d3 = rhoR * wR - rhoL * wL;
  const double d4 = ER - EL;
  const double dp = gm1 * (d4 - uh * d1 - vh * d2 - wh * d3 + 0.5 * q2 * d0);
  const double dun = (d1 * n[0] + d2 * n[1] + d3 * n[2]) - unh * d0;
  const double rcd = rhoh * ch * dun;
  const double a1 = (dp - rcd) * (0.5 * rc2), a5 = (dp + rcd) * (0.5 * rc2);
  const double a23 = d0 - dp * rc2;
  const double rrh = rcp_hd(rhoh);
  const double e1 = d1 - uh * d0, e2 = d2 - vh * d0, e3 = d3 - wh * d0;
  const double dut = e1 * rrh, dvt = e2 * rrh, dwt = e3 * rrh;
  const double dunf = dut * n[0] + dvt * n[1] + dwt * n[2];
  const double tx = dut - dunf * n[0], ty = dvt - dunf * n[1], tz = dwt - dunf * n[2];
  const double unL = uL * n[0] + vL * n[1] + wL * n[2], unR = uR * n[0] + vR * n[1] + wR * n[2];
  const double w1 = al1 * a1, w5 = al5 * a5, w2 = al2 * a23, w2r = al2 * rhoh;
  const double uht = uh * tx + vh * ty + wh * tz;

  // ---------------- reverse ----------------
  const double f0 = lF[0], f1 = lF[1], f2 = lF[2], f3 = lF[3], f4 = lF[4];
  double b_rhoL = 0, b_rhoR = 0, b_uL = 0, b_vL = 0, b_wL = 0, b_pL = 0, b_uR = 0, b_vR = 0, b_wR = 0, b_pR = 0;
  double b_unL = 0, b_unR = 0, b_EL = 0, b_ER = 0;
  double b_w1 = 0, b_w5 = 0, b_w2 = 0, b_w2r = 0, b_uh = 0, b_vh = 0, b_wh = 0, b_Hh = 0, b_q2 = 0;
  double b_ch = 0, b_unh = 0, b_tx = 0, b_ty = 0, b_tz = 0;
  // F0
  b_rhoL += 0.5 * f0 * unL; b_unL += 0.5 * f0 * rhoL; b_rhoR += 0.5 * f0 * unR; b_unR += 0.5 * f0 * rhoR;
  b_w1 -= 0.5 * f0; b_w5 -= 0.5 * f0; b_w2 -= 0.5 * f0;
  // F1..F3
  const double fm[3] = {f1, f2, f3};
  const double uLv[3] = {uL, vL, wL}, uRv[3] = {uR, vR, wR}, uhv[3] = {uh, vh, wh}, tv[3] = {tx, ty, tz};
  double b_uLv[3] = {0, 0, 0}, b_uRv[3] = {0, 0, 0}, b_uhv[3] = {0, 0, 0}, b_tv[3] = {0, 0, 0};
  for (int m = 0; m < 3; ++m) {
    const double fv = fm[m];
    const double cn = ch * n[m];
    b_rhoL += 0.5 * fv * uLv[m] * unL; b_uLv[m] += 0.5 * fv * rhoL * unL; b_unL += 0.5 * fv * rhoL * uLv[m];
    b_pL += 0.5 * fv * n[m];
    b_rhoR += 0.5 * fv * uRv[m] * unR; b_uRv[m] += 0.5 * fv * rhoR * unR; b_unR += 0.5 * fv * rhoR * uRv[m];
    b_pR += 0.5 * fv * n[m];
    b_w1 -= 0.5 * fv * (uhv[m] - cn); b_w5 -= 0.5 * fv * (uhv[m] + cn); b_w2 -= 0.5 * fv * uhv[m];
    b_w2r -= 0.5 * fv * tv[m];
    b_uhv[m] -= 0.5 * fv * (w1 + w5 + w2);
    b_ch -= 0.5 * fv * (w5 - w1) * n[m];          // through cn = ch n_m
    b_tv[m] -= 0.5 * fv * w2r;
  }
  // F4
  b_EL += 0.5 * f4 * unL; b_pL += 0.5 * f4 * unL; b_unL += 0.5 * f4 * (EL + pL);
  b_ER += 0.5 * f4 * unR; b_pR += 0.5 * f4 * unR; b_unR += 0.5 * f4 * (ER + pR);
  {
    const double cun = ch * unh;
    b_w1 -= 0.5 * f4 * (Hh - cun); b_w5 -= 0.5 * f4 * (Hh + cun);
    b_w2 -= 0.5 * f4 * (0.5 * q2); b_w2r -= 0.5 * f4 * uht;
    b_Hh -= 0.5 * f4 * (w1 + w5);
    const double b_cun = -0.5 * f4 * (w5 - w1);
    b_ch += b_cun * unh; b_unh += b_cun * ch;
    b_q2 -= 0.5 * f4 * w2 * 0.5;
    for (int m = 0; m < 3; ++m) {
      b_uhv[m] -= 0.5 * f4 * w2r * tv[m];
      b_tv[m] -= 0.5 * f4 * w2r * uhv[m];
    }
  }
  b_uh += b_uhv[0]; b_vh += b_uhv[1]; b_wh += b_uhv[2];
  b_tx += b_tv[0]; b_ty += b_tv[1]; b_tz += b_tv[2];
  b_uL += b_uLv[0]; b_vL += b_uLv[1]; b_wL += b_uLv[2];
  b_uR += b_uRv[0]; b_vR += b_uRv[1]; b_wR += b_uRv[2];
  // w1 = al1 a1, w5 = al5 a5, w2 = al2 a23, w2r = al2 rhoh
  double b_al1 = b_w1 * a1, b_a1 = b_w1 * al1, b_al5 = b_w5 * a5, b_a5 = b_w5 * al5;
  double b_al2 = b_w2 * a23 + b_w2r * rhoh, b_a23 = b_w2 * al2, b_rhoh = b_w2r * al2;
  // unL, unR
  b_uL += b_unL * n[0]; b_vL += b_unL * n[1]; b_wL += b_unL * n[2];
  b_uR += b_unR * n[0]; b_vR += b_unR * n[1]; b_wR += b_unR * n[2];
  // t = dv - dunf n;  dunf = dv . n
  const double b_dunf = -(b_tx * n[0] + b_ty * n[1] + b_tz * n[2]);
  const double b_dut = b_tx + b_dunf * n[0], b_dvt = b_ty + b_dunf * n[1], b_dwt = b_tz + b_dunf * n[2];
  // dut = (d1 - uh d0) rrh etc.
  double b_d0 = 0, b_d1 = 0, b_d2 = 0, b_d3 = 0, b_d4 = 0;
  b_d1 += b_dut * rrh; b_uh -= b_dut * d0 * rrh; b_d0 -= b_dut * uh * rrh;
  b_d2 += b_dvt * rrh; b_vh -= b_dvt * d0 * rrh; b_d0 -= b_dvt * vh * rrh;
  b_d3 += b_dwt * rrh; b_wh -= b_dwt * d0 * rrh; b_d0 -= b_dwt * wh * rrh;
  const double b_rrh = b_dut * e1 + b_dvt * e2 + b_dwt * e3;
  b_rhoh -= b_rrh * rrh * rrh;
  // a23 = d0 - dp rc2 ; a1 = (dp - rcd) rc2/2 ; a5 = (dp + rcd) rc2/2
  double b_dp = 0, b_rc2 = 0, b_rcd = 0;
  b_d0 += b_a23; b_dp -= b_a23 * rc2; b_rc2 -= b_a23 * dp;
  b_dp += b_a1 * 0.5 * rc2; b_rcd -= b_a1 * 0.5 * rc2; b_rc2 += b_a1 * 0.5 * (dp - rcd);
  b_dp += b_a5 * 0.5 * rc2; b_rcd += b_a5 * 0.5 * rc2; b_rc2 += b_a5 * 0.5 * (dp + rcd);
  // rcd = rhoh ch dun
  b_rhoh += b_rcd * ch * dun; b_ch += b_rcd * rhoh * dun;
  const double b_dun = b_rcd * rhoh * ch;
  // dun = d.n - unh d0
  b_d1 += b_dun * n[0]; b_d2 += b_dun * n[1]; b_d3 += b_dun * n[2]; b_unh -= b_dun * d0; b_d0 -= b_dun * unh;
  // dp = gm1 (d4 - uh d1 - vh d2 - wh d3 + q2/2 d0)
  {
    const double e = gm1 * b_dp;
    b_d4 += e; b_uh -= e * d1; b_d1 -= e * uh; b_vh -= e * d2; b_d2 -= e * vh; b_wh -= e * d3; b_d3 -= e * wh;
    b_q2 += e * 0.5 * d0; b_d0 += e * 0.5 * q2;
  }
  // d0..d4
  b_rhoR += b_d0; b_rhoL -= b_d0;
  b_rhoR += b_d1 * uR + b_d2 * vR + b_d3 * wR;
  b_rhoL -= b_d1 * uL + b_d2 * vL + b_d3 * wL;
  b_uR += b_d1 * rhoR; b_vR += b_d2 * rhoR; b_wR += b_d3 * rhoR;
  b_uL -= b_d1 * rhoL; b_vL -= b_d2 * rhoL; b_wL -= b_d3 * rhoL;
  b_ER += b_d4; b_EL -= b_d4;
  // al2 = |unh|
  b_unh += b_al2 * dabs(unh);
  // Harten-fixed al1, al5
  double b_eps = 0.0, b_lam1 = 0.0, b_lam5 = 0.0;
  if (br1) b_lam1 += b_al1 * dabs(lam1);
  else { b_lam1 += b_al1 * lam1 * reps; b_eps += b_al1 * 0.5 * (1.0 - lam1 * lam1 * reps * reps); }
  if (br5) b_lam5 += b_al5 * dabs(lam5);
  else { b_lam5 += b_al5 * lam5 * reps; b_eps += b_al5 * 0.5 * (1.0 - lam5 * lam5 * reps * reps); }
  b_unh += b_lam1 + b_lam5; b_ch += b_lam5 - b_lam1;
  // eps = 0.1 (|unh| + ch)
  b_unh += 0.1 * b_eps * dabs(unh); b_ch += 0.1 * b_eps;
  // rhoh = sqL sqR
  double b_sqL = b_rhoh * sqR, b_sqR = b_rhoh * sqL;
  // unh = uh.n
  b_uh += b_unh * n[0]; b_vh += b_unh * n[1]; b_wh += b_unh * n[2];
  // ch = sqrt(c2), rc2 = 1/c2, c2 = max(c2raw, 1e-30)
  double b_c2 = b_ch * 0.5 * (ch * rc2) - b_rc2 * rc2 * rc2;   // 1/ch = ch/c2
  const double b_c2raw = b_c2 * dmax_floor(c2raw, 1e-30);
  b_Hh += gm1 * b_c2raw; b_q2 -= 0.5 * gm1 * b_c2raw;
  // q2 = |uh|^2
  b_uh += 2.0 * uh * b_q2; b_vh += 2.0 * vh * b_q2; b_wh += 2.0 * wh * b_q2;
  // uh = nuh rs, ... Hh = nHh rs
  double b_rs = b_uh * nuh + b_vh * nvh + b_wh * nwh + b_Hh * nHh;
  b_sqL += (b_uh * uL + b_vh * vL + b_wh * wL + b_Hh * HL) * rs;
  b_sqR += (b_uh * uR + b_vh * vR + b_wh * wR + b_Hh * HR) * rs;
  b_uL += b_uh * sqL * rs; b_vL += b_vh * sqL * rs; b_wL += b_wh * sqL * rs;
  b_uR += b_uh * sqR * rs; b_vR += b_vh * sqR * rs; b_wR += b_wh * sqR * rs;
  const double b_HL = b_Hh * sqL * rs, b_HR = b_Hh * sqR * rs;
  // rs = 1/max(S, 1e-150)
  const double b_S = -b_rs * rs * rs * dmax_floor(S, 1e-150);
  b_sqL += b_S; b_sqR += b_S;
  // sq = sqrt(rho)
  b_rhoL += b_sqL * 0.5 * (sqL * irL); b_rhoR += b_sqR * 0.5 * (sqR * irR);   // 1/sqrt(rho) = sqrt(rho)/rho
  // H = (E + p)/rho
  b_EL += b_HL * irL; b_pL += b_HL * irL; b_rhoL -= b_HL * HL * irL;
  b_ER += b_HR * irR; b_pR += b_HR * irR; b_rhoR -= b_HR * HR * irR;
  // E = p/(g-1) + rho ke
  b_pL += b_EL * rgm1; b_rhoL += b_EL * keL;
  b_pR += b_ER * rgm1; b_rhoR += b_ER * keR;
  const double b_keL = b_EL * rhoL, b_keR = b_ER * rhoR;
  b_uL += b_keL * uL; b_vL += b_keL * vL; b_wL += b_keL * wL;
  b_uR += b_keR * uR; b_vR += b_keR * vR; b_wR += b_keR * wR;
  // rho = max(pv0, 1e-300)
  lpl[0] += b_rhoL * dmax_floor(pl[0], 1e-300);
  lpr[0] += b_rhoR * dmax_floor(pr[0], 1e-300);
  lpl[1] += b_uL; lpl[2] += b_vL; lpl[3] += b_wL; lpl[4] += b_pL;
  lpr[1] += b_uR; lpr[2] += b_vR; lpr[3] += b_wR; lpr[4] += b_pR;
}

// Reverse of one inviscid face (face_flux_prim): given lFa = cotangent of
// F * |n| for the face between cells a and a+1 whose left/right states come
// from cells a-1..a+2, return the cotangents of the reconstructed states and
// of the two limited slopes:  pL = pv1 + s1/2, pR = pv2 - s2/2.
template <int LIM>
FD void face_states(const double pv0[5], const double pv1[5], const double pv2[5], const double pv3[5],
                    double pL[5], double pR[5]) {
  for (int v = 0; v < 5; ++v) {
    const double sL = limit(LIM, pv1[v] - pv0[v], pv2[v] - pv1[v]);
    const double sR = limit(LIM, pv2[v] - pv1[v], pv3[v] - pv2[v]);
    pL[v] = pv1[v] + 0.5 * sL;
    pR[v] = pv2[v] - 0.5 * sR;
  }
}

FD void face_normal(const double k0[3], const double k1[3], double nu[3], double& nm) {
  const double nf0 = 0.5 * (k0[0] + k1[0]), nf1 = 0.5 * (k0[1] + k1[1]), nf2 = 0.5 * (k0[2] + k1[2]);
  nm = sqrt_hd(nf0 * nf0 + nf1 * nf1 + nf2 * nf2) + 1e-300;
  const double rn = rcp_hd(nm);
  nu[0] = nf0 * rn; nu[1] = nf1 * rn; nu[2] = nf2 * rn;
}

// Viscous flux tensor (rhs/viscous.py:30-108) reverse.  Inputs as
// viscous_flux; lG = cotangent of G[9]; outputs accumulate the cotangents of
// u, v, w (direct use in the energy flux), T (through mu and kappa) and the
// central differences d[l][s].
FD void viscous_flux_vjp(const double pv[5], double T, const double d[3][4], const double k[9], double gamma,
                         double R, double mu0, double T0, double Ts, double Pr, const double lG[9], double& lu,
                         double& lv, double& lw, double& lT, double ld[3][4]) {
  const double x = T * rcp_hd(T0);
  const double sx = sqrt_hd(x);
  const double rTs = rcp_hd(T + Ts);
  const double mu = mu0 * (x * sx) * (T0 + Ts) * rTs;
  const double cp_pr = gamma * R * rcp_hd(gamma - 1.0) * rcp_hd(Pr);
  const double kappa = cp_pr * mu;
  double gr[4][3];
  for (int s = 0; s < 4; ++s)
    for (int m = 0; m < 3; ++m) gr[s][m] = d[0][s] * k[0 * 3 + m] + d[1][s] * k[1 * 3 + m] + d[2][s] * k[2 * 3 + m];
  const double divu3 = (gr[0][0] + gr[1][1] + gr[2][2]) * (1.0 / 3.0);
  const double t00 = 2 * mu * (gr[0][0] - divu3), t01 = mu * (gr[0][1] + gr[1][0]), t02 = mu * (gr[0][2] + gr[2][0]);
  const double t11 = 2 * mu * (gr[1][1] - divu3), t12 = mu * (gr[1][2] + gr[2][1]), t22 = 2 * mu * (gr[2][2] - divu3);
  const double u = pv[1], v = pv[2], w = pv[3];
  // E_j = u t0j + v t1j + w t2j + kappa dT/dx_j
  const double b00 = lG[0] + u * lG[6];
  const double b01 = lG[1] + v * lG[6] + u * lG[7];
  const double b02 = lG[2] + w * lG[6] + u * lG[8];
  const double b11 = lG[3] + v * lG[7];
  const double b12 = lG[4] + w * lG[7] + v * lG[8];
  const double b22 = lG[5] + w * lG[8];
  lu += t00 * lG[6] + t01 * lG[7] + t02 * lG[8];
  lv += t01 * lG[6] + t11 * lG[7] + t12 * lG[8];
  lw += t02 * lG[6] + t12 * lG[7] + t22 * lG[8];
  const double b_kappa = gr[3][0] * lG[6] + gr[3][1] * lG[7] + gr[3][2] * lG[8];
  double bg[4][3];
  bg[3][0] = kappa * lG[6]; bg[3][1] = kappa * lG[7]; bg[3][2] = kappa * lG[8];
  const double b_mu = 2.0 * (gr[0][0] - divu3) * b00 + (gr[0][1] + gr[1][0]) * b01 + (gr[0][2] + gr[2][0]) * b02 +
                      2.0 * (gr[1][1] - divu3) * b11 + (gr[1][2] + gr[2][1]) * b12 +
                      2.0 * (gr[2][2] - divu3) * b22 + cp_pr * b_kappa;
  const double b_div3 = -2.0 * mu * (b00 + b11 + b22);
  bg[0][0] = 2.0 * mu * b00 + b_div3 * (1.0 / 3.0);
  bg[1][1] = 2.0 * mu * b11 + b_div3 * (1.0 / 3.0);
  bg[2][2] = 2.0 * mu * b22 + b_div3 * (1.0 / 3.0);
  bg[0][1] = mu * b01; bg[1][0] = mu * b01;
  bg[0][2] = mu * b02; bg[2][0] = mu * b02;
  bg[1][2] = mu * b12; bg[2][1] = mu * b12;
  for (int l = 0; l < 3; ++l)
    for (int s = 0; s < 4; ++s)
      ld[l][s] += bg[s][0] * k[l * 3 + 0] + bg[s][1] * k[l * 3 + 1] + bg[s][2] * k[l * 3 + 2];
  // mu = mu0 x^1.5 (T0+Ts)/(T+Ts):  dmu/dT = mu (1.5/T - 1/(T+Ts))
  lT += b_mu * mu * (1.5 * rcp_hd(T) - rTs);
}

// T = p / max(rho R, 1e-300) reverse into (rho, p).
FD void temperature_vjp(const double pv[5], double R, double lT, double& lrho, double& lp) {
  const double den = pv[0] * R;
  const double dn = max_floor(den, 1e-300);
  const double rd = rcp_hd(dn);
  lp += lT * rd;
  lrho -= lT * pv[4] * rd * rd * R * dmax_floor(den, 1e-300);
}

// No-slip isothermal ghost value (physics/bc/wall.py:35-58) reverse: lo =
// cotangent of the ghost cell, accumulated into li (the mirrored inner cell).
FD void isothermal_vjp(const double qi[5], double gamma, double R, double Tw, const double lo[5], double li[5]) {
  const double gm1 = gamma - 1.0;
  const double rho_i = qi[0];
  const double ssum = qi[1] * qi[1] + qi[2] * qi[2] + qi[3] * qi[3];
  const double ke_i = 0.5 * ssum / rho_i;
  const double p = gm1 * (qi[4] - ke_i);
  const double rho_g = p / (R * Tw);
  const double scale = rho_g / rho_i;
  const double m[3] = {-qi[1] * scale, -qi[2] * scale, -qi[3] * scale};
  const double ke_g = 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rho_g;
  double b_rho_g = lo[0], b_p = lo[4] / gm1, b_keg = lo[4];
  double b_m[3] = {lo[1], lo[2], lo[3]};
  for (int k = 0; k < 3; ++k) b_m[k] += b_keg * m[k] / rho_g;
  b_rho_g -= b_keg * ke_g / rho_g;
  double b_scale = 0.0;
  for (int k = 0; k < 3; ++k) { li[1 + k] -= b_m[k] * scale; b_scale -= b_m[k] * qi[1 + k]; }
  b_rho_g += b_scale / rho_i;
  double b_rho_i = -b_scale * scale / rho_i;
  b_p += b_rho_g / (R * Tw);
  li[4] += gm1 * b_p;
  const double b_kei = -gm1 * b_p;
  for (int k = 0; k < 3; ++k) li[1 + k] += b_kei * qi[1 + k] / rho_i;
  b_rho_i -= b_kei * ke_i / rho_i;
  li[0] += b_rho_i;
}

}  // namespace famrdp
