// Adjoint (reverse-mode) kernels around the stage reverse (famrdp_reverse.hip):
// the BC faces and the halo copy reversed (BC faces in reverse Face order, the
// halo through a source-sorted table), the objectives and their gradients, and
// the state copies of the checkpointing driver.  Functions cite the forward
// code they reverse in famrdp_jax (physics/bc/*.py, mesh/halo.py).
#include "famrdp_adjoint.h"

namespace famrdp {

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

// Deterministic sum of the farfield partials of each face entry, in two
// passes: BC_RED_PARTS workgroups per entry (blockIdx.y) each sum a fixed
// chunk of the face's cells into red[(entry * 5 + v) * BC_RED_PARTS + part],
// then k_bc_reduce_final adds the parts of each entry in order to
// dqinf[(block*6 + face)*5 + v].  One workgroup per face left the reduction
// of a 10^4-cell face to 256 threads in series.
constexpr int BC_RED_PARTS = 32;

__global__ void __launch_bounds__(256) k_bc_reduce(const BlockDesc* __restrict__ blocks,
                                                   const BcEntry* __restrict__ entries, const double* __restrict__ part,
                                                   int64_t maxcells, double* __restrict__ red) {
  const BcEntry e = entries[blockIdx.y];
  if (e.kind != BC_FARFIELD) return;
  const BlockDesc b = blocks[e.block];
  const int ax = e.face / 2;
  const int n[3] = {b.ni, b.nj, b.nk};
  const int64_t ncell = (int64_t)n[ax == 0 ? 1 : 0] * n[ax == 2 ? 1 : 2];
  const int64_t chunk = (ncell + BC_RED_PARTS - 1) / BC_RED_PARTS;
  const int64_t t0 = blockIdx.x * chunk, t1 = t0 + chunk < ncell ? t0 + chunk : ncell;
  __shared__ double acc[5][256];
  for (int v = 0; v < 5; ++v) {
    double s = 0.0;
    for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) s += part[((int64_t)blockIdx.y * 5 + v) * maxcells + t];
    acc[v][threadIdx.x] = s;
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int v = 0; v < 5; ++v) acc[v][threadIdx.x] += acc[v][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 5) red[((int64_t)blockIdx.y * 5 + threadIdx.x) * BC_RED_PARTS + blockIdx.x] = acc[threadIdx.x][0];
}

__global__ void k_bc_reduce_final(const BcEntry* __restrict__ entries, const double* __restrict__ red,
                                  double* __restrict__ dqinf) {
  const BcEntry e = entries[blockIdx.x];
  const int v = threadIdx.x;
  if (e.kind != BC_FARFIELD || v >= 5 || !dqinf) return;
  double s = 0.0;
  for (int q = 0; q < BC_RED_PARTS; ++q) s += red[((int64_t)blockIdx.x * 5 + v) * BC_RED_PARTS + q];
  dqinf[(e.block * 6 + e.face) * 5 + v] += s;
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
