/* C ABI of the MI355X implementation of famrdp_jax's RHS + TVD-RK3 path.
 *
 * Plain C types only: device pointers to float64 arrays in the reference's
 * layout, sizes, and an opaque stream handle (a hipStream_t passed as void*;
 * NULL = the default stream).  Every function returns 0 on success or a
 * negative FAMRDP_E_* code; famrdp_last_error() gives the message of the
 * last failure on the calling thread.  No function allocates device memory
 * except famrdp_plan_create (which sizes the plan's workspace once).
 *
 * Array layout (famrdp_jax/core/types.py:26-93): q (5, ni, nj, nk), kxyz
 * (3, 3, ni, nj, nk) with kxyz[l][m] = d xi_l / d x_m, vol and jac
 * (ni, nj, nk), xyz (3, ni, nj, nk); C order, k unit stride; ni, nj, nk are
 * padded sizes (interior + 2*ghost).
 *
 * Citations are to the reference repository (LucasLu123/famrdp-jax).
 */
#ifndef FAMRDP_CAPI_H
#define FAMRDP_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FAMRDP_OK 0
#define FAMRDP_E_INVALID (-1)      /* bad argument (Python: ValueError / TypeError)        */
#define FAMRDP_E_HIP (-2)          /* HIP runtime error                                     */
#define FAMRDP_E_UNSUPPORTED (-3)  /* e.g. orient != 0 (Python: NotImplementedError)       */
#define FAMRDP_E_KEY (-4)          /* unknown limiter name (Python: KeyError)               */

/* Limiters: famrdp_jax/physics/reconstruct/muscl2.py:15 (_LIMITERS). */
#define FAMRDP_LIMITER_MINMOD 0
#define FAMRDP_LIMITER_VANLEER 1

/* Ghost-cell BC kinds (famrdp_jax/physics/bc/dispatch.py:11-32). */
#define FAMRDP_BC_WALL_ADIABATIC 0   /* bc/wall.py:30-32                       */
#define FAMRDP_BC_WALL_ISOTHERMAL 1  /* bc/wall.py:35-58                       */
#define FAMRDP_BC_MIRROR_NORMAL 2    /* bc/symmetry.py:8-11, bc/wall.py:61-64  */
#define FAMRDP_BC_FARFIELD 3         /* bc/farfield.py:7-21                    */
#define FAMRDP_BC_EXTRAPOLATE 4      /* extension: not in the reference        */

/* famrdp_jax/core/types.py:158-167 (GasModel). */
typedef struct famrdp_gas {
  double gamma, R, mu0, T0, Ts, Pr_lam;
} famrdp_gas;

/* One CUT1TO1 copy of famrdp_jax/mesh/halo.py:79-90, listed in the order the
 * reference visits them (blocks in order, faces in neighbors-dict order).
 * Faces are famrdp_jax/core/constants.py:55-61 (IMIN=0 .. KMAX=5). */
typedef struct famrdp_cut {
  int32_t tgt_block, tgt_face, src_block, src_face, orient;
} famrdp_cut;

/* One BC face (block index, face, FAMRDP_BC_* kind, T_wall, q_inf). */
typedef struct famrdp_bc {
  int32_t block, face, kind, pad_;
  double T_wall;
  double q_inf[5];
} famrdp_bc;

typedef struct famrdp_plan famrdp_plan;

/* Message of the last error on this thread ("" if none). */
const char* famrdp_last_error(void);
/* Library / kernel variant identification (static string). */
const char* famrdp_version(void);

/* ---- pointwise and operator entry points -------------------------------- */

/* famrdp_jax/physics/eos.py:12-17. q, pv: (5, n). */
int famrdp_cons_to_prim(const double* q, double* pv, int64_t n, double gamma, void* stream);
/* famrdp_jax/physics/eos.py:5-9. pv, q: (5, n). */
int famrdp_prim_to_cons(const double* pv, double* q, int64_t n, double gamma, void* stream);
/* famrdp_jax/physics/eos.py:24-25. */
int famrdp_sutherland(const double* T, double* mu, int64_t n, double mu0, double T0, double Ts, void* stream);
/* famrdp_jax/physics/eos.py:20-21. */
int famrdp_speed_of_sound(const double* rho, const double* p, double* c, int64_t n, double gamma, void* stream);
/* famrdp_jax/physics/reconstruct/muscl2.py:5-12: out = minmod(a, b) or
 * vanleer(a, b) elementwise. */
int famrdp_limiter(const double* a, const double* b, double* out, int64_t n, int32_t limiter, void* stream);
/* famrdp_jax/physics/reconstruct/muscl2.py:18-37.  pv (nvar, n0, n1, n2);
 * axis in 1..3; pv_L/pv_R have n_axis-3 along the axis. */
int famrdp_muscl2(const double* pv, double* pv_L, double* pv_R, int32_t nvar, int32_t n0, int32_t n1, int32_t n2,
                  int32_t axis, int32_t limiter, void* stream);
/* famrdp_jax/flux/roe.py:21-85.  UL, UR, F: (5, n); nxyz (3, n), unit normals. */
int famrdp_roe_flux(const double* UL, const double* UR, const double* nxyz, double* F, int64_t n, double gamma,
                    double eps_entropy, void* stream);
/* famrdp_jax/rhs/invscd.py:18-133 (reconstruct/flux names are unused there). */
int famrdp_invscd_rhs(const double* q, const double* kxyz, const double* vol, double* rhs, int32_t ni, int32_t nj,
                      int32_t nk, double gamma, int32_t limiter, void* stream);
/* famrdp_jax/rhs/viscous.py:19-119.  work: caller scratch of 9*ni*nj*nk doubles. */
int famrdp_viscous_rhs(const double* q, const double* kxyz, double* rhs, int32_t ni, int32_t nj, int32_t nk,
                       const famrdp_gas* gas, double* work, void* stream);
/* famrdp_jax/time/rk.py:17-19 combination: out = a*x + b*(y + c*z), or
 * y + c*z when mode == 1 (x may be NULL then). */
int famrdp_axpby(double* out, double a, const double* x, double b, const double* y, double c, const double* z,
                 int64_t n, int32_t mode, void* stream);
/* famrdp_jax/mesh/metric.py:31-70.  xyz (3, ni, nj, nk) -> jac, kxyz, vol. */
int famrdp_compute_metrics(const double* xyz, double* jac, double* kxyz, double* vol, int32_t ni, int32_t nj,
                           int32_t nk, void* stream);

/* The stage kernels' reciprocal, square root and reciprocal square root
 * (hardware seed + Newton refinement) on n values, for verifying their
 * accuracy against IEEE 1/x, sqrt(x) and 1/sqrt(x) (out_rsqrt may be NULL).
 * No counterpart in the reference. */
int famrdp_fastmath_probe(const double* x, double* out_rcp, double* out_sqrt, double* out_rsqrt, int64_t n,
                          void* stream);
/* The face flux's reciprocal and reciprocal square root (hardware seed + one
 * Newton step) on n values, for verifying their accuracy against IEEE 1/x
 * and 1/sqrt(x) (x > 0).  No counterpart in the reference. */
int famrdp_fastmath_probe1(const double* x, double* out_rcp, double* out_rsqrt, int64_t n, void* stream);

/* ---- plan: one multi-block configuration (Config.topology + gas + scheme) - */

/* dims: nblocks x {ni, nj, nk} (padded, each >= 2*ghost; a block of 2^31 or
 * more cells, or of 2^29 or more with FAMRDP_PLAN_ADJOINT, is refused with
 * FAMRDP_E_INVALID).  cuts/bcs as above.  Builds
 * the halo gather table (final writer per ghost cell, pre-exchange sources),
 * the per-face BC tables, the tile table and the stage workspace.
 * Replaces famrdp_jax/solver/step.py:18-56's per-call setup. */
#define FAMRDP_PLAN_NO_WORKSPACE 1  /* halo / BC only: no stage buffers (famrdp_step is refused) */
#define FAMRDP_PLAN_ADJOINT 2       /* also size the reverse-sweep workspace (famrdp_step_vjp) */
int famrdp_plan_create(famrdp_plan** out, int32_t nblocks, const int32_t* dims, int32_t ncut, const famrdp_cut* cuts,
                       int32_t nbc, const famrdp_bc* bcs, const famrdp_gas* gas, int32_t limiter, int32_t ghost,
                       int32_t plan_flags);
int famrdp_plan_destroy(famrdp_plan* plan);

/* Step variants (flags of famrdp_step). */
#define FAMRDP_STEP_BATCHED 1   /* one launch per kernel across all blocks (block-offset/tile table) */
#define FAMRDP_STEP_DIRECT 2    /* reference-order direct kernels (one thread per cell, no LDS); for cross-checks */
#define FAMRDP_VJP_CONSUME_START 4  /* famrdp_step_vjp*: q_start may be overwritten (it receives the halo and BCs
                                       the reverse sweep linearises at) instead of being copied first */

/* famrdp_jax/solver/step.py:18-56: halo (mesh/halo.py:56), BCs
 * (physics/bc/dispatch.py:11), then Shu-Osher RK3 on every padded cell,
 * in place on q[b].  Pointer arrays have nblocks entries (host memory). */
int famrdp_step(famrdp_plan* plan, double* const* q, const double* const* kxyz, const double* const* vol, double dt,
                int32_t flags, void* stream);
/* nsteps steps replayed from a hipGraph captured for these pointers, dt and
 * flags (captured on first use, re-captured when any of them changes). */
int famrdp_step_graph(famrdp_plan* plan, double* const* q, const double* const* kxyz, const double* const* vol,
                      double dt, int32_t flags, int32_t nsteps, void* stream);
/* famrdp_step that writes the step's stage-1 and stage-2 states (q1 and q2
 * of solver/step.py:45-53, locals there) into q1_out / q2_out (per block,
 * device 5*N doubles, distinct from q and from each other) instead of the
 * plan's workspace: the states famrdp_step_vjp_stages takes, without a copy.
 * Same step and bits as famrdp_step; famrdp_plan_stage_states keeps
 * returning the workspace of the last famrdp_step / famrdp_step_graph. */
int famrdp_step_stages(famrdp_plan* plan, double* const* q, const double* const* kxyz, const double* const* vol,
                       double dt, int32_t flags, double* const* q1_out, double* const* q2_out, void* stream);
/* Stage-kernel selection (no counterpart in the reference).  Variants 0-4
 * and 10-13 are tile shapes of the fused stage kernel (4 and 10-13 stage the
 * primitives twice to need less LDS), 5-9 tile shapes of the split pair
 * (viscous tensor kernel + stage kernel), 14-18 shapes of the i-marching
 * kernel (a TJ x TK column walks a segment of TI planes along i; same bits as
 * the fused kernel; blocks under 59,652,323 cells, else set_tile returns
 * FAMRDP_E_INVALID and the autotuner reports 1e30 ms and skips them).  famrdp_plan_autotune times `reps`
 * stage-1 + stage-2 launch pairs of every variant for these arrays (both write
 * only the plan's workspace, so the state is not modified; stage 2 also reads
 * q0 like stage 3), writes the per-variant ms of one pair to ms_out (may be
 * NULL; FAMRDP_NUM_TILE_VARIANTS entries) and keeps the fastest; a variant
 * within 1.5 % of an earlier one counts as a tie and the earlier one stays. */
#define FAMRDP_NUM_TILE_VARIANTS 19
int famrdp_plan_autotune(famrdp_plan* plan, double* const* q, const double* const* kxyz, const double* const* vol,
                         double dt, int32_t flags, int32_t reps, double* ms_out, void* stream);
int famrdp_plan_set_tile(famrdp_plan* plan, int32_t variant);
/* variant index and its tile dims {TI, TJ, TK} (cells per workgroup per axis). */
int famrdp_plan_get_tile(const famrdp_plan* plan, int32_t* variant, int32_t* dims3);
/* Stage timing for measurement: enable=1 starts recording a HIP event pair
 * around every RK3 stage's kernels of eager famrdp_step calls (on the launch
 * stream); enable=0 stops, synchronizes and returns the summed stage time (ms)
 * and the number of stages timed.  No counterpart in the reference (its
 * benchmarks/bench_step.py:35-51 times whole steps on the host). */
int famrdp_plan_timing(famrdp_plan* plan, int32_t enable, double* stage_ms, int64_t* stage_count);
/* Per-block Sutherland mu0 and time-step scale (block b advances by
 * dt * dt_scale[b]) for ensembles of independent cases in one plan; either
 * array may be NULL.  No counterpart in the reference (one GasModel per
 * Config, core/types.py:158-189). */
int famrdp_plan_set_block_params(famrdp_plan* plan, const double* mu0, const double* dt_scale);

/* Reverse of one famrdp_step (solver/step.py:18-56 under jax.grad, see
 * tests/solver/test_grad.py:10-38): given q_start (the state the step started
 * from, per block) and lam (cotangent of the step's result, per block, device
 * 5*N doubles), overwrites lam with the cotangent of q_start (lam is the
 * sweep's accumulator from the start: the caller must not read it before the
 * stream reaches the end of the call).  Recomputes
 * halo, BCs and stages 1-2, reverses the three stages (viscous, Roe, MUSCL),
 * the BCs in reverse Face order and the halo copy.  dqinf (device, NULL or
 * nblocks*6*5 doubles) accumulates the cotangent of each (block, face)
 * farfield state q_inf.  flags as famrdp_step (the recompute uses the same
 * kernels).  Needs a plan created with FAMRDP_PLAN_ADJOINT. */
int famrdp_step_vjp(famrdp_plan* plan, double* const* q_start, const double* const* kxyz, const double* const* vol,
                    double* const* lam, double dt, double* dqinf, int32_t flags, void* stream);
/* famrdp_step_vjp with the step's stage-1 and stage-2 states supplied (per
 * block, device 5*N doubles, as written by famrdp_step_stages or saved by
 * famrdp_plan_stage_states right after the forward famrdp_step from q_start)
 * instead of recomputed.  Same reverse
 * sweep and results; no counterpart in the reference (checkpointed reverse
 * passes that keep a segment's stage states). */
int famrdp_step_vjp_stages(famrdp_plan* plan, double* const* q_start, const double* const* q1, const double* const* q2,
                           const double* const* kxyz, const double* const* vol, double* const* lam, double dt,
                           double* dqinf, int32_t flags, void* stream);
/* Copies the stage-1 and stage-2 states of the plan's last famrdp_step or
 * famrdp_step_graph (its workspace) into q1_out / q2_out (per block, device
 * 5*N doubles).  No counterpart in the reference. */
int famrdp_plan_stage_states(famrdp_plan* plan, double* const* q1_out, double* const* q2_out, void* stream);

/* Build-defined objectives for the differentiation workloads (the reference
 * defines none): kind 0 = interior kinetic energy sum 1/2 rho|u|^2 vol,
 * kind 1 = integrated KMIN wall pressure sum_{i,j} p |grad zeta| vol at
 * k = ghost.  J (device, 1 double); grad (device, 5*N) = dJ/dq on every cell;
 * work (device) >= 1024 doubles. */
int famrdp_objective(int32_t kind, const double* q, const double* kxyz, const double* vol, int32_t ni, int32_t nj,
                     int32_t nk, int32_t ghost, double gamma, double* J, double* grad, double* work, void* stream);

/* famrdp_jax/mesh/halo.py:56-92 alone, in place. */
int famrdp_halo_exchange(famrdp_plan* plan, double* const* q, void* stream);
/* famrdp_jax/physics/bc/dispatch.py:11-32 alone, in place. */
int famrdp_apply_bc(famrdp_plan* plan, double* const* q, void* stream);

/* Build-defined CFL time step over interior cells of one block plus the
 * positivity monitor (Config.cfl / Config.nan_check, core/types.py:187-189,
 * unused in the reference).  out (device, 3 doubles) = {dt, max spectral
 * radius, number of non-finite / non-positive cells}.  work: 2*1024 doubles. */
int famrdp_cfl_dt(const double* q, const double* kxyz, const double* vol, int32_t ni, int32_t nj, int32_t nk,
                  int32_t ghost, const famrdp_gas* gas, double cfl, double* out, double* work, void* stream);

/* No counterpart in the reference: famrdp_cfl_dt without the work argument,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-workgroup partial results go to a library-owned device array of
 * 2 x 1024 doubles (16 KiB on each GPU the library runs on, held until the
 * process ends); its size does not depend on the block, so every block
 * famrdp_cfl_dt accepts fits.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two threads) race.  Calls no
 * allocating routine; same results as famrdp_cfl_dt. */
int famrdp_cfl_dt_noalloc(const double* q, const double* kxyz, const double* vol, int32_t ni, int32_t nj,
                          int32_t nk, int32_t ghost, const famrdp_gas* gas, double cfl, double* out, void* stream);

/* Config.nan_check (famrdp_jax/core/types.py:187, carried but never read by
 * the reference): out (device, 1 double) = number of interior cells, over
 * every block of the plan, with a non-finite value, rho <= 0 or p <= 0.
 * Stream-ordered; the caller reads out when it needs the answer. */
int famrdp_plan_check_state(famrdp_plan* plan, double* const* q, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FAMRDP_CAPI_H */
