"""GPU adjoint (famrdp_step_vjp + checkpointing) against the autograd oracle
(torch float64 through the oracle step, JAX derivative conventions) and
against central finite differences (h = 1e-6).  Bar: 1e-8 relative."""
import numpy as np
import pytest
import torch

from famrdp_jax import cases
from famrdp_jax.adjoint import OBJ_KINETIC_ENERGY, OBJ_WALL_PRESSURE, Adjoint, mach_derivative, steps_autograd
from famrdp_jax.core.constants import BCType, Face
from famrdp_jax.core.types import Block, BlockTopology, Config, GasModel, Metrics, SchemeChoice, State
from oracle import blocks as OB
from oracle import metric as OM
from _util import to_dev
from _adjoint_util import _c4_small, _oracle_grad

pytestmark = pytest.mark.gpu
TOL = 1e-8


def _rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _where(a, b):
    """Where the largest gradient errors sit: per variable, per i/j/k index."""
    d = np.abs(np.asarray(a) - np.asarray(b)) / max(np.abs(b).max(), 1e-300)
    idx = np.unravel_index(np.argmax(d), d.shape)
    by = {ax: np.round(d.max(axis=tuple(x for x in range(4) if x != ax)), 12).tolist() for ax in (0, 1, 2, 3)}
    return f"argmax {idx}; by var {by[0]}; by i {by[1]}; by j {by[2]}; by k {by[3]}"


@pytest.mark.parametrize("kind", [OBJ_KINETIC_ENERGY, OBJ_WALL_PRESSURE])
def test_wall_farfield_gradient_vs_autograd(kind):
    print(f"case: C4-like 12^3 wall/farfield, 3 steps, objective {kind}", flush=True)
    state, cfg, dt = _c4_small(12)
    adj = Adjoint(to_dev(state), cfg)
    J, g, dqinf = adj.gradient(dt, 3, kind, checkpoint_every=2)
    Jr, gr, qr = _oracle_grad(state, cfg, dt, 3, kind)
    e = _rel(g[0].cpu().numpy(), gr[0])
    eq = _rel(dqinf[0, int(Face.KMAX)].cpu().numpy(), qr[int(Face.KMAX)])
    print(f"  J {J:.12e} vs {Jr:.12e}; dJ/dq0 rel {e:.2e}; dJ/dq_inf rel {eq:.2e}", flush=True)
    if e > TOL:
        print("  " + _where(g[0].cpu().numpy(), gr[0]), flush=True)
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    assert e <= TOL and eq <= TOL


@pytest.mark.parametrize("limiter", ["minmod", "vanleer"])
def test_periodic_gradient_vs_autograd(limiter):
    print(f"case: periodic 10^3 {limiter}, 4 steps, checkpoint 3", flush=True)
    st, cfg = cases.case_c1(n=10, seed=7, p_noise=1e-3)
    cfg = Config(gas=cfg.gas, scheme=SchemeChoice("muscl2", "roe", 3, limiter), bc_params={}, topology=cfg.topology,
                 ghost=2)
    m = st.metrics[0]
    dt = OM.cfl_dt(st.blocks[0].q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    adj = Adjoint(to_dev(st), cfg)
    J, g, _ = adj.gradient(dt, 4, OBJ_KINETIC_ENERGY, checkpoint_every=3)
    Jr, gr, _ = _oracle_grad(st, cfg, dt, 4, OBJ_KINETIC_ENERGY)
    e = _rel(g[0].cpu().numpy(), gr[0])
    print(f"  dJ/dq0 rel {e:.2e}", flush=True)
    if e > TOL:
        print("  " + _where(g[0].cpu().numpy(), gr[0]), flush=True)
    assert e <= TOL


def test_multiblock_batched_gradient_vs_autograd():
    print("case: 2x2x2 blocks (cut/farfield/extrapolate), 2 steps, batched", flush=True)
    st, cfg = cases.case_c3(nb_axis=2, lo=6, hi=9, seed=11, p_glue=0.7, p_noise=1e-3)
    dt = min(OM.cfl_dt(b.q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5) for b, m in zip(st.blocks, st.metrics))
    adj = Adjoint(to_dev(st), cfg, batched=True)
    J, g, dqinf = adj.gradient(dt, 2, OBJ_KINETIC_ENERGY, checkpoint_every=10)
    Jr, gr, qr = _oracle_grad(st, cfg, dt, 2, OBJ_KINETIC_ENERGY)
    e = max(_rel(a.cpu().numpy(), b) for a, b in zip(g, gr))
    # the farfield q_inf is shared by every FARFIELD face (bc_params keyed by face)
    eq = _rel(dqinf[:, int(Face.IMIN)].sum(0).cpu().numpy(), qr[int(Face.IMIN)])
    print(f"  dJ/dq0 rel {e:.2e}; dJ/dq_inf rel {eq:.2e}", flush=True)
    assert e <= TOL and eq <= TOL


def test_ensemble_gradient_per_block_mu0_and_dt():
    """C5-like ensemble in one batched adjoint plan (own mu0 and dt per
    block): J = sum of the cases' kinetic energies, gradient per case against
    the autograd oracle of that case alone."""
    print("case: C5-like ensemble of 3 cases, 3 steps, batched adjoint", flush=True)
    pairs = cases.case_c5(ncases=3, seed=8, sizes=(8, 10, 12), p_noise=1e-3)
    state, cfg, mu0 = cases.ensemble(pairs)
    dts = [OM.cfl_dt(st.blocks[0].q, st.metrics[0].kxyz, st.metrics[0].vol, cases.gas_dict(c.gas), 0.5)
           for st, c in pairs]
    adj = Adjoint(to_dev(state), cfg, batched=True)
    adj.set_block_params(mu0, dts)
    J, g, _ = adj.gradient(1.0, 3, OBJ_KINETIC_ENERGY, checkpoint_every=2)
    Jr, errs = 0.0, []
    for c, ((st, cc), dt) in enumerate(zip(pairs, dts)):
        Jc, gc, _ = _oracle_grad(st, cc, dt, 3, OBJ_KINETIC_ENERGY)
        Jr += Jc
        errs.append(_rel(g[c].cpu().numpy(), gc[0]))
    print(f"  J {J:.12e} vs {Jr:.12e}; dJ/dq0 rel per case {[f'{e:.2e}' for e in errs]}", flush=True)
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    assert max(errs) <= TOL


def test_reference_grad_case_finite_nonzero_and_matches():
    """reference tests/solver/test_grad.py:10-38: uniform farfield flow,
    J = sum(q after one step); gradient finite, non-zero, and equal to the
    autograd oracle (v = w = 0 puts u.n = 0 exactly on j/k faces)."""
    print("case: reference test_grad 12x10x8", flush=True)
    g = 1.4
    rho, u, p = 1.225, 50.0, 101325.0
    qinf = np.array([rho, rho * u, 0.0, 0.0, p / (g - 1) + 0.5 * rho * u * u])
    q = np.broadcast_to(qinf[:, None, None, None], (5, 12, 10, 8)).copy()
    m = cases.cartesian_metrics((12, 10, 8))
    topo = BlockTopology(0, {f: None for f in Face}, {f: BCType.FARFIELD for f in Face})
    cfg = Config(gas=GasModel(), scheme=SchemeChoice("muscl2", "roe", 3, "minmod"),
                 bc_params={f: {"q_inf": qinf} for f in Face}, topology=(topo,), ghost=2, nan_check=False)
    st = State(blocks=(Block(q=q, xyz=np.zeros((3, 12, 10, 8))),), metrics=(m,))
    adj = Adjoint(to_dev(st), cfg)
    lam = [torch.ones_like(adj.q0[0])]
    adj.reverse(lam, 1e-6, 1, {0: [adj.q0[0].clone()]}, 10, None)
    torch.cuda.synchronize()
    gq = lam[0].cpu().numpy()
    assert np.isfinite(gq).all() and (gq != 0).any()
    _, gr, _ = _oracle_grad(st, cfg, 1e-6, 1, None, seed_all=True)
    print(f"  grad max {np.abs(gq).max():.4e}; rel vs oracle {_rel(gq, gr[0]):.2e}", flush=True)
    if _rel(gq, gr[0]) > TOL:
        print("  " + _where(gq, gr[0]), flush=True)
    assert _rel(gq, gr[0]) <= TOL


def test_central_finite_differences_field_and_mach():
    """Directional derivative of the wall-pressure J along a field direction,
    and dJ/dM_inf of the kinetic-energy J through the farfield state (after 4
    steps the farfield barely reaches the wall: the wall J's Mach derivative,
    ~5e-7, is below what h = 1e-6 resolves), against central differences
    with h = 1e-6, at the 1e-8 bar.

    The step is piecewise smooth: the minmod limiter, |u_n| and the Harten
    fix switch branches wherever a slope or a wave speed crosses a kink, and
    this noisy state has such kinks within ~1e-7 relative of many cells, so a
    random field direction crosses some of them at h = 1e-6 and no
    derivative matches its difference quotient.  The field direction is
    therefore the gradient's own (scaled so the largest change of a cell is
    1e-7 of max|q|), which turns within the smooth piece around q0.  Two
    roundoff floors are printed: the worst-case eps |J| / (h |dJ|) and the
    measured one, the change of the difference quotient when h grows by
    2^-10 (truncation changes by ~h^2 2^-9 there, roundoff does not
    correlate).  The field direction is scaled so that the measured floor is
    <= 1e-9; the Mach quotient's floor (set by |J| / |dJ/dM| at h = 1e-6)
    must stay below the 1e-8 bar."""
    print("case: C4-like 12^3 FD checks (h = 1e-6)", flush=True)
    state, cfg, dt = _c4_small(12)
    nsteps = 4
    dev = to_dev(state)
    adj = Adjoint(dev, cfg)
    h = 1e-6
    eps = float(np.finfo(np.float64).eps)

    def J_of(kind, q0, qinf=None):
        c = cfg if qinf is None else Config(gas=cfg.gas, scheme=cfg.scheme, bc_params={
            Face.KMIN: cfg.bc_params[Face.KMIN], Face.KMAX: {"q_inf": qinf}}, topology=cfg.topology, ghost=2)
        a = Adjoint(State(blocks=(Block(q=q0, xyz=dev.blocks[0].xyz),), metrics=dev.metrics), c)
        out = a.forward(dt, nsteps)
        return a.objective(out, kind)[0]

    J0, g, _ = adj.gradient(dt, nsteps, OBJ_WALL_PRESSURE, checkpoint_every=2)
    v = g[0] / g[0].abs().max() * adj.q0[0].abs().max() * 0.05
    ad = float((g[0] * v).sum())
    fds = [(J_of(OBJ_WALL_PRESSURE, adj.q0[0] + hh * v) - J_of(OBJ_WALL_PRESSURE, adj.q0[0] - hh * v)) / (2 * hh)
           for hh in (h, h * (1 + 2.0 ** -10))]
    rel, noise, worst = abs(fds[0] - ad) / abs(ad), abs(fds[0] - fds[1]) / abs(ad), eps * abs(J0) / (h * abs(ad))
    print(f"  field: FD {fds[0]:.12e} adjoint {ad:.12e} rel {rel:.2e} (floor measured {noise:.2e}, "
          f"worst case {worst:.2e})", flush=True)
    assert noise <= 1e-9
    assert rel <= 1e-8
    J1, _, dqinf = adj.gradient(dt, nsteps, OBJ_KINETIC_ENERGY, checkpoint_every=2)
    M = 0.8
    fdM = []
    for hh in (h, h * (1 + 2.0 ** -10)):
        Jp = J_of(OBJ_KINETIC_ENERGY, adj.q0[0], cases.freestream(M + hh))
        Jm = J_of(OBJ_KINETIC_ENERGY, adj.q0[0], cases.freestream(M - hh))
        fdM.append((Jp - Jm) / (2 * hh))
    adM = mach_derivative(dqinf, M)
    rel, noise, worst = abs(fdM[0] - adM) / abs(adM), abs(fdM[0] - fdM[1]) / abs(adM), eps * abs(J1) / (h * abs(adM))
    print(f"  Mach: FD {fdM[0]:.12e} adjoint {adM:.12e} rel {rel:.2e} (floor measured {noise:.2e}, "
          f"worst case {worst:.2e})", flush=True)
    # A scalar parameter has no direction to rescale: at h = 1e-6 this
    # quotient's roundoff is set by |J| / |dJ/dM| (measured ~6e-9), so the
    # check asserts that the floor stays below the 1e-8 bar it guards.
    assert noise <= 1e-8
    assert rel <= 1e-8


def test_kept_stage_states_equal_recomputed():
    """The reverse pass that keeps each step's stage-1/2 states
    (famrdp_plan_stage_states + famrdp_step_vjp_stages) gives the same bits as
    the one that recomputes them (famrdp_step_vjp)."""
    print("case: C4-like 12^3, 3 steps, kept vs recomputed stage states", flush=True)
    state, cfg, dt = _c4_small(12)
    adj = Adjoint(to_dev(state), cfg)
    final, ckpts = adj.forward_checkpointed(adj.q0, dt, 3, 2)
    J, lam0 = adj.objective(final, OBJ_WALL_PRESSURE)
    outs = []
    for keep in (True, False):
        lam = [l.clone() for l in lam0]
        dq = torch.zeros(6 * 5, dtype=torch.float64, device="cuda")
        adj.reverse(lam, dt, 3, ckpts, 2, dq, keep_stages=keep)
        torch.cuda.synchronize()
        outs.append((lam[0].cpu().numpy(), dq.cpu().numpy()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    # every step's states kept from the forward pass (no recompute): same bits
    J0, g0, dq0 = adj.gradient(dt, 3, OBJ_WALL_PRESSURE, checkpoint_every=0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(g0[0].cpu().numpy(), outs[0][0])
    np.testing.assert_array_equal(dq0.reshape(-1).cpu().numpy(), outs[0][1])
    # checkpointed, the last segment's states kept by the forward pass: same bits
    J2, g2, dq2 = adj.gradient(dt, 3, OBJ_WALL_PRESSURE, checkpoint_every=2)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(g2[0].cpu().numpy(), outs[0][0])
    np.testing.assert_array_equal(dq2.reshape(-1).cpu().numpy(), outs[0][1])


@pytest.mark.parametrize("batched", [False, True])
def test_step_vjp_consumed_or_copied_start_same_bits(batched):
    """One famrdp_step_vjp four ways -- q_start copied or consumed
    (FAMRDP_VJP_CONSUME_START), stage states recomputed or supplied -- gives
    the same cotangent and q_inf cotangent bit for bit; without the flag
    q_start is left as it was, with it only its ghost cells change (the halo
    and BCs the reverse sweep linearises at)."""
    print(f"case: C4-like 12^3, one reverse step, batched={batched}, copied/consumed start", flush=True)
    state, cfg, dt = _c4_small(12)
    adj = Adjoint(to_dev(state), cfg, batched=batched)
    start = [q.clone() for q in adj.q0]
    cur = [q.clone() for q in adj.q0]
    adj._step(cur, dt)
    q1s = [torch.empty_like(q) for q in cur]
    q2s = [torch.empty_like(q) for q in cur]
    adj.plan.stage_states(q1s, q2s)
    _, lam0 = adj.objective(cur, OBJ_WALL_PRESSURE)
    outs = []
    for consume in (False, True):
        for stages in (None, (q1s, q2s)):
            qs = [q.clone() for q in start]
            lam = [l.clone() for l in lam0]
            dq = torch.zeros(6 * 5, dtype=torch.float64, device="cuda")
            adj.plan.step_vjp(qs, adj.kxyz, adj.vol, lam, dt, dq, batched=batched, stages=stages,
                              consume_start=consume)
            torch.cuda.synchronize()
            if not consume:
                assert torch.equal(qs[0], start[0])
            else:
                g = cfg.ghost
                assert torch.equal(qs[0][:, g:-g, g:-g, g:-g], start[0][:, g:-g, g:-g, g:-g])
            outs.append((lam[0].cpu().numpy(), dq.cpu().numpy()))
    for lam_o, dq_o in outs[1:]:
        np.testing.assert_array_equal(lam_o, outs[0][0])
        np.testing.assert_array_equal(dq_o, outs[0][1])
    assert np.isfinite(outs[0][0]).all() and np.abs(outs[0][0]).max() > 0


@pytest.mark.parametrize("batched,direct", [(False, False), (True, False), (False, True)])
def test_step_stages_writes_the_stage_states(batched, direct):
    """famrdp_step_stages gives the same step, bit for bit, as famrdp_step and
    writes into the caller's arrays the stage states famrdp_plan_stage_states
    copies out after famrdp_step; a famrdp_step after it (back on the plan's
    workspace) is unchanged too, and stage_states still returns its states."""
    print(f"case: C4-like 12^3, famrdp_step_stages vs famrdp_step, batched={batched}, direct={direct}", flush=True)
    state, cfg, dt = _c4_small(12)
    adj = Adjoint(to_dev(state), cfg, batched=batched, direct=direct)
    a = [q.clone() for q in adj.q0]
    adj._step(a, dt)
    r1 = [torch.empty_like(q) for q in a]
    r2 = [torch.empty_like(q) for q in a]
    adj.plan.stage_states(r1, r2)
    b = [q.clone() for q in adj.q0]
    q1s, q2s = adj._step_stages(b, dt)
    torch.cuda.synchronize()
    for x, y in ((a, b), (r1, q1s), (r2, q2s)):
        assert torch.equal(x[0], y[0])
    assert not torch.equal(q1s[0], q2s[0]) and not torch.equal(q1s[0], b[0])
    # the next plain step runs on the workspace again: same bits both ways
    adj._step(a, dt)
    adj._step(b, dt)
    s1 = [torch.empty_like(q) for q in a]
    s2 = [torch.empty_like(q) for q in a]
    adj.plan.stage_states(s1, s2)
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0])
    assert not torch.equal(s1[0], q1s[0]) and torch.isfinite(s1[0]).all()
    # a captured graph around it keeps stepping on the workspace: graph, stages, graph
    # steps give the bits of three plain steps
    c = [q.clone() for q in adj.q0]
    d = [q.clone() for q in adj.q0]
    for _ in range(3):
        adj._step(c, dt)
    adj.plan.step(d, adj.kxyz, adj.vol, dt, batched=batched, direct=direct, graph_steps=1)
    adj._step_stages(d, dt)
    adj.plan.step(d, adj.kxyz, adj.vol, dt, batched=batched, direct=direct, graph_steps=1)
    torch.cuda.synchronize()
    assert torch.equal(c[0], d[0])
    # the stage arrays must be distinct from the state and from each other
    with pytest.raises(ValueError):
        adj.plan.step(b, adj.kxyz, adj.vol, dt, batched=batched, direct=direct, stages=(b, q2s))
    with pytest.raises(ValueError):
        adj.plan.step(b, adj.kxyz, adj.vol, dt, batched=batched, direct=direct, stages=(q1s, q1s))


def test_torch_autograd_function():
    print("case: torch.autograd.Function wrapper", flush=True)
    st, cfg = cases.case_c1(n=8, seed=9, p_noise=1e-3)
    m = st.metrics[0]
    dt = OM.cfl_dt(st.blocks[0].q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    adj = Adjoint(to_dev(st), cfg)
    q0 = adj.q0[0].clone().requires_grad_(True)
    out = steps_autograd(q0, adj, dt, 3, 2)
    (out ** 2).sum().backward()
    qt = torch.tensor(st.blocks[0].q, requires_grad=True)
    cur = [qt]
    for _ in range(3):
        cur = OB.step(cur, [(torch.tensor(m.kxyz), torch.tensor(m.vol))], list(cfg.topology), {},
                      cases.gas_dict(cfg.gas), dt)
    (cur[0] ** 2).sum().backward()
    assert _rel(q0.grad.cpu().numpy(), qt.grad.numpy()) <= TOL
