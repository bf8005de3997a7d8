// Test-only host build of the hand-written reverse sweeps in
// famrdp-jax_amd/csrc/famrdp_adjoint.h, so the CPU tests can compare them with
// autograd through the oracle.  Never loaded by the product.
#include "famrdp_adjoint.h"

using namespace famrdp;

extern "C" {

// n samples; arrays are sample-major ([n][5] etc.).
void hc_roe_vjp(int n, const double* pl, const double* pr, const double* nrm, double g, const double* lF, double* lpl,
                double* lpr) {
  for (int i = 0; i < n; ++i) {
    for (int v = 0; v < 5; ++v) { lpl[5 * i + v] = 0.0; lpr[5 * i + v] = 0.0; }
    roe_flux_prim_vjp(pl + 5 * i, pr + 5 * i, nrm + 3 * i, g, lF + 5 * i, lpl + 5 * i, lpr + 5 * i);
  }
}

void hc_cons_to_prim_vjp(int n, const double* q, double g, const double* lp, double* lq) {
  for (int i = 0; i < n; ++i) {
    for (int v = 0; v < 5; ++v) lq[5 * i + v] = 0.0;
    cons_to_prim_vjp(q + 5 * i, g, lp + 5 * i, lq + 5 * i);
  }
}

// the stage kernels' minmod (famrdp_math.h minmod_fast)
void hc_minmod_fast(int n, const double* a, const double* b, double* out) {
  for (int i = 0; i < n; ++i) out[i] = minmod_fast(a[i], b[i]);
}

void hc_limiter_vjp(int n, int lim, const double* A, const double* B, const double* ls, double* lA, double* lB) {
  for (int i = 0; i < n; ++i) limiter_vjp(lim, A[i], B[i], ls[i], lA[i], lB[i]);
}

// pv [n][5], T [n], d [n][3][4], k [n][9], lG [n][9] -> out [n][4 + 12]
void hc_viscous_vjp(int n, const double* pv, const double* T, const double* d, const double* k, const double* gas,
                    const double* lG, double* out) {
  for (int i = 0; i < n; ++i) {
    double lu = 0, lv = 0, lw = 0, lT = 0, ld[3][4] = {};
    double dd[3][4];
    for (int l = 0; l < 3; ++l)
      for (int s = 0; s < 4; ++s) dd[l][s] = d[12 * i + 4 * l + s];
    viscous_flux_vjp(pv + 5 * i, T[i], dd, k + 9 * i, gas[0], gas[1], gas[2], gas[3], gas[4], gas[5], lG + 9 * i, lu,
                     lv, lw, lT, ld);
    out[16 * i + 0] = lu; out[16 * i + 1] = lv; out[16 * i + 2] = lw; out[16 * i + 3] = lT;
    for (int l = 0; l < 3; ++l)
      for (int s = 0; s < 4; ++s) out[16 * i + 4 + 4 * l + s] = ld[l][s];
  }
}

void hc_isothermal_vjp(int n, const double* qi, double g, double R, double Tw, const double* lo, double* li) {
  for (int i = 0; i < n; ++i) {
    for (int v = 0; v < 5; ++v) li[5 * i + v] = 0.0;
    isothermal_vjp(qi + 5 * i, g, R, Tw, lo + 5 * i, li + 5 * i);
  }
}
}
