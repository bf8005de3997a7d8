"""GPU parity of the full TVD-RK3 step (halo + BC + 3 stages) against the
oracle, at the sizes the oracle runs in seconds, plus size-independent
properties at the full 256^3 size.

Bars (famrdp_jax semantics, FP64): one step from the same state <= 1e-12
relative L-inf; free-running divergence after 100 steps <= 1e-9.
"""
import numpy as np
import pytest
import torch

from famrdp_jax import _lib, cases
from famrdp_jax.solver.step import step
from famrdp_jax.solver.stepper import Stepper
from oracle import blocks as OB
from oracle import metric as OM
from _util import rel_linf, to_dev

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
LONG_TOL = 1e-9


def _oracle_args(state, cfg):
    return ([m.kxyz for m in state.metrics], [m.vol for m in state.metrics])


def _ostep(qs, state, cfg, dt):
    mets = [(m.kxyz, m.vol) for m in state.metrics]
    return OB.step(qs, mets, list(cfg.topology), cfg.bc_params, cases.gas_dict(cfg.gas), dt)


def _gpu_step_from(qs, state, cfg, dt, **kw):
    from famrdp_jax.core.types import Block, State
    st = State(blocks=tuple(Block(q=q, xyz=b.xyz) for q, b in zip(qs, state.blocks)), metrics=state.metrics)
    out = step(to_dev(st), cfg, dt)
    return [b.q.cpu().numpy() for b in out.blocks]


def test_c1_ten_steps_stepwise_and_free_running():
    print("case: C1 32^3 periodic, 10 steps", flush=True)
    state, cfg = cases.case_c1(n=32, seed=0)
    m = state.metrics[0]
    dt = OM.cfl_dt(state.blocks[0].q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    ref = [state.blocks[0].q]
    st = Stepper(to_dev(state), cfg)
    worst = 0.0
    for n in range(10):
        nxt = _ostep(ref, state, cfg, dt)
        one = _gpu_step_from(ref, state, cfg, dt)
        worst = max(worst, rel_linf(one[0], nxt[0]))
        ref = nxt
        st.advance(dt)
    torch.cuda.synchronize()
    free = rel_linf(st.qs[0].cpu().numpy(), ref[0])
    print(f"C1 worst per-step rel Linf {worst:.3e}, after 10 steps {free:.3e}", flush=True)
    assert worst <= STEP_TOL
    assert free <= LONG_TOL


def _c2_small(n):
    q, xyz_ext, cfg = cases.case_c2(n)
    jac, k, vol = OM.compute_metrics(xyz_ext)
    m = cases.crop_metrics(jac, k, vol)
    from famrdp_jax.core.types import Block, State
    state = State(blocks=(Block(q=q, xyz=cases.crop_xyz(xyz_ext)),), metrics=(m,))
    dt = OM.cfl_dt(q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    return state, cfg, dt


def test_c2_grid_curvilinear_100_steps():
    """Config 2's grid and TGV at 24^3 (oracle-sized), 100 steps."""
    print("case: C2 grid 24^3 curvilinear TGV, 100 steps", flush=True)
    state, cfg, dt = _c2_small(24)
    ref = [state.blocks[0].q]
    st = Stepper(to_dev(state), cfg)
    worst = 0.0
    for n in range(100):
        nxt = _ostep(ref, state, cfg, dt)
        if n % 10 == 0:
            worst = max(worst, rel_linf(_gpu_step_from(ref, state, cfg, dt)[0], nxt[0]))
        ref = nxt
    st.advance(dt, 100, graph=True)
    torch.cuda.synchronize()
    free = rel_linf(st.qs[0].cpu().numpy(), ref[0])
    print(f"C2-24 worst per-step {worst:.3e}, after 100 steps {free:.3e}", flush=True)
    assert worst <= STEP_TOL
    assert free <= LONG_TOL


def _small_multiblock():
    return cases.case_c3(nb_axis=2, lo=8, hi=14, seed=11, p_glue=0.7)


def test_multiblock_plain_batched_graph():
    print("case: C3-like 2x2x2 blocks, 3 steps (plain / batched / graph)", flush=True)
    state, cfg = _small_multiblock()
    dt = min(OM.cfl_dt(b.q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5) for b, m in zip(state.blocks, state.metrics))
    ref = [b.q for b in state.blocks]
    for _ in range(3):
        ref = _ostep(ref, state, cfg, dt)
    outs = {}
    for mode in ("plain", "batched", "graph", "graph_batched"):
        st = Stepper(to_dev(state), cfg)
        st.advance(dt, 3, batched="batched" in mode, graph="graph" in mode)
        torch.cuda.synchronize()
        outs[mode] = [q.cpu().numpy() for q in st.qs]
        err = max(rel_linf(a, b) for a, b in zip(outs[mode], ref))
        print(f"  {mode}: rel Linf after 3 steps {err:.3e}", flush=True)
        assert err <= 1e-11
    for mode in ("batched", "graph", "graph_batched"):
        for a, b in zip(outs[mode], outs["plain"]):
            np.testing.assert_array_equal(a, b)


def test_direct_kernels_match_oracle_and_tiled():
    """The reference-order direct kernels (second, independent GPU
    implementation) against the oracle and the tiled kernels, on the C1
    workload and on an uneven multi-block case."""
    print("case: direct vs tiled vs oracle (C1 16^3, 2x2x2 blocks)", flush=True)
    state, cfg = cases.case_c1(n=16, seed=4)
    m = state.metrics[0]
    dt = OM.cfl_dt(state.blocks[0].q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    ref = _ostep([state.blocks[0].q], state, cfg, dt)[0]
    for direct in (True, False):
        st = Stepper(to_dev(state), cfg)
        st.advance(dt, direct=direct)
        torch.cuda.synchronize()
        assert rel_linf(st.qs[0].cpu().numpy(), ref) <= STEP_TOL
    state, cfg = _small_multiblock()
    dt = min(OM.cfl_dt(b.q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5) for b, m in zip(state.blocks, state.metrics))
    outs = []
    for kw in (dict(direct=True), dict(direct=True, batched=True), dict(), dict(batched=True)):
        st = Stepper(to_dev(state), cfg)
        st.advance(dt, 2, **kw)
        torch.cuda.synchronize()
        outs.append([q.cpu().numpy() for q in st.qs])
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(outs[2], outs[3]):
        np.testing.assert_array_equal(a, b)
    assert max(rel_linf(a, b) for a, b in zip(outs[2], outs[0])) <= 1e-12


def test_tile_variants_bitwise_identical_and_autotune():
    """Every tile shape of one kernel form computes each face and viscous
    tensor with the same inputs and operation order, so the variants of the
    fused kernel give bit-identical states, and so do those of the split pair;
    the two forms differ only in where the compiler contracts a product into
    an FMA (the split pair stores T before differencing it), <= 1e-13.  The
    autotuner leaves the state untouched."""
    print("case: tile variants (C1 20^3 and 2x2x2 blocks)", flush=True)
    for state, cfg in (cases.case_c1(n=20, seed=5), _small_multiblock()):
        dt = 0.5 * min(OM.cfl_dt(b.q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
                       for b, m in zip(state.blocks, state.metrics))
        ref = {}
        for v in range(_lib.NUM_TILE_VARIANTS):
            st = Stepper(to_dev(state), cfg)
            st.plan.set_tile(v)
            st.advance(dt, 2, batched=(v % 2 == 1))
            torch.cuda.synchronize()
            out = [q.cpu().numpy() for q in st.qs]
            form = "fused" if v in _lib.FUSED_VARIANTS else "split"
            if form not in ref:
                ref[form] = out
            else:
                for a, b in zip(out, ref[form]):
                    np.testing.assert_array_equal(a, b)
        err = max(rel_linf(a, b) for a, b in zip(ref["split"], ref["fused"]))
        print(f"  split vs fused rel Linf {err:.3e}", flush=True)
        assert err <= 1e-13
        st = Stepper(to_dev(state), cfg)
        before = [q.clone() for q in st.qs]
        v, dims, ms = st.autotune(dt, reps=1)
        assert 0 <= v < _lib.NUM_TILE_VARIANTS and len(ms) == _lib.NUM_TILE_VARIANTS and all(x > 0 for x in ms)
        for a, b in zip(st.qs, before):
            assert torch.equal(a, b)
        st.plan.set_tile(0)


@pytest.mark.parametrize("n", [13, 14, 40])
def test_tile_variants_single_wrap_blocks_vs_oracle(n):
    """Blocks whose padded sizes are all >= 18 take the tile kernels' single
    conditional wrap, smaller ones a modulo: the C2 grid at 13^3 (17^3
    padded: modulo), 14^3 (18^3: the smallest single-wrap size, where a
    16-cell tile's halo reaches 15 cells past the block end) and 40^3 (44^3:
    a partial last tile for every tile shape; the 8 x 32 marching variants
    take the 12-column k remainder in 32 x 8 strip tiles), periodic wrap at both ends of
    each axis; one step per variant against the oracle, fused variants
    bitwise among themselves."""
    print(f"case: C2 grid {n}^3, every tile variant vs oracle, 1 step", flush=True)
    state, cfg, dt = _c2_small(n)
    ref = _ostep([state.blocks[0].q], state, cfg, dt)[0]
    fused = None
    for v in range(_lib.NUM_TILE_VARIANTS):
        st = Stepper(to_dev(state), cfg)
        st.plan.set_tile(v)
        st.advance(dt, batched=(v % 2 == 1))
        torch.cuda.synchronize()
        out = st.qs[0].cpu().numpy()
        err = rel_linf(out, ref)
        print(f"  variant {v}: rel Linf {err:.3e}", flush=True)
        assert err <= STEP_TOL
        if v in _lib.FUSED_VARIANTS:
            if fused is None:
                fused = out
            else:
                np.testing.assert_array_equal(out, fused)


def test_tile_variants_vanleer_vs_oracle():
    """The van Leer limiter (a template instance of every stage kernel of its
    own): every variant, plain and batched, one step against the oracle on
    the C2 grid at 14^3 (18^3 padded, single wrap) and 40^3 (44^3: strip
    tiles of the marching variants) and on the small uneven multi-block case
    (modulo wrap); fused variants bitwise among themselves."""
    from famrdp_jax.core.types import Config, SchemeChoice
    print("case: van Leer, every tile variant vs oracle (C2 grid 14^3 and 40^3, 2x2x2 blocks), 1 step", flush=True)
    for state, cfg in (_c2_small(14)[:2], _c2_small(40)[:2], _small_multiblock()):
        cfg = Config(gas=cfg.gas, scheme=SchemeChoice("muscl2", "roe", 3, "vanleer"), bc_params=cfg.bc_params,
                     topology=cfg.topology, ghost=2)
        dt = 0.5 * min(OM.cfl_dt(b.q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
                       for b, m in zip(state.blocks, state.metrics))
        mets = [(m.kxyz, m.vol) for m in state.metrics]
        ref = OB.step([b.q for b in state.blocks], mets, list(cfg.topology), cfg.bc_params, cases.gas_dict(cfg.gas),
                      dt, limiter="vanleer")
        fused = None
        for v in range(_lib.NUM_TILE_VARIANTS):
            for batched in (False, True):
                st = Stepper(to_dev(state), cfg)
                st.plan.set_tile(v)
                st.advance(dt, batched=batched)
                torch.cuda.synchronize()
                out = [q.cpu().numpy() for q in st.qs]
                err = max(rel_linf(a, b) for a, b in zip(out, ref))
                assert err <= STEP_TOL, (v, batched, err)
                if v in _lib.FUSED_VARIANTS:
                    if fused is None:
                        fused = out
                    else:
                        for a, b in zip(out, fused):
                            np.testing.assert_array_equal(a, b)
        print(f"  {len(state.blocks)} block(s): every variant within {STEP_TOL:g}", flush=True)


@pytest.mark.parametrize("n", [1, 2])
def test_tile_variants_degenerate_blocks_vs_oracle(n):
    """Blocks with one or two interior cells (5^3 and 6^3 padded): every tile
    of every variant overhangs the block several times over along each axis
    (wrapped by the modulo form), no cell or only one gets the inviscid
    update; one step per variant against the oracle."""
    print(f"case: C1 {n}^3 interior, every tile variant vs oracle, 1 step", flush=True)
    state, cfg = cases.case_c1(n=n, seed=3)
    m = state.metrics[0]
    dt = OM.cfl_dt(state.blocks[0].q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    ref = _ostep([state.blocks[0].q], state, cfg, dt)[0]
    for v in range(_lib.NUM_TILE_VARIANTS):
        st = Stepper(to_dev(state), cfg)
        st.plan.set_tile(v)
        st.advance(dt, batched=(v % 2 == 1))
        torch.cuda.synchronize()
        err = rel_linf(st.qs[0].cpu().numpy(), ref)
        print(f"  variant {v}: rel Linf {err:.3e}", flush=True)
        assert err <= STEP_TOL


def test_c5_ensemble_per_block_mu0_and_dt():
    """Independent periodic cases with their own Mach, Reynolds number (mu0)
    and CFL time step, advanced together by one batched plan
    (set_block_params, dt = 1, dt_scale = dt_case), against the oracle run
    case by case; the per-block plain and direct paths agree with it."""
    print("case: C5-like ensemble of 3 cases (sizes 8-12), 3 steps", flush=True)
    pairs = cases.case_c5(ncases=3, seed=8, sizes=(8, 10, 12))
    state, cfg, mu0 = cases.ensemble(pairs)
    dts = [OM.cfl_dt(st.blocks[0].q, st.metrics[0].kxyz, st.metrics[0].vol, cases.gas_dict(c.gas), 0.5)
           for st, c in pairs]
    assert len(set(mu0)) == 3 and len(set(dts)) == 3
    refs = []
    for (st, c), dt in zip(pairs, dts):
        q = [st.blocks[0].q]
        for _ in range(3):
            q = _ostep(q, st, c, dt)
        refs.append(q[0])
    outs = {}
    for name, kw in (("batched", dict(batched=True)), ("plain", dict()), ("graph_batched", dict(batched=True, graph=True)),
                     ("direct_batched", dict(batched=True, direct=True))):
        stp = Stepper(to_dev(state), cfg, mu0=mu0, dt_scale=dts)
        stp.advance(1.0, 3, **kw)
        torch.cuda.synchronize()
        outs[name] = [q.cpu().numpy() for q in stp.qs]
        err = max(rel_linf(a, b) for a, b in zip(outs[name], refs))
        print(f"  {name}: rel Linf after 3 steps {err:.3e}", flush=True)
        assert err <= 1e-11
    for name in ("plain", "graph_batched"):
        for a, b in zip(outs[name], outs["batched"]):
            np.testing.assert_array_equal(a, b)


def test_c4_wall_farfield_small():
    print("case: C4-like 16^3 wall/farfield, 5 steps", flush=True)
    q, xyz_ext, cfg = cases.case_c4(n=16)
    jac, k, vol = OM.compute_metrics(xyz_ext)
    m = cases.crop_metrics(jac, k, vol)
    from famrdp_jax.core.types import Block, State
    state = State(blocks=(Block(q=q, xyz=cases.crop_xyz(xyz_ext)),), metrics=(m,))
    dt = OM.cfl_dt(q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    ref = [q]
    worst = 0.0
    for _ in range(5):
        nxt = _ostep(ref, state, cfg, dt)
        worst = max(worst, rel_linf(_gpu_step_from(ref, state, cfg, dt)[0], nxt[0]))
        ref = nxt
    print(f"C4-16 worst per-step {worst:.3e}", flush=True)
    assert worst <= STEP_TOL


def test_reference_uniform_flow_farfield():
    """reference tests/solver/test_step.py:7-42 through the GPU step."""
    print("case: uniform farfield 16x12x10, 3 steps", flush=True)
    from famrdp_jax.core.constants import BCType, Face
    from famrdp_jax.core.types import Block, BlockTopology, Config, GasModel, Metrics, SchemeChoice, State
    g = 1.4
    rho, u, p = 1.225, 50.0, 101325.0
    qinf = np.array([rho, rho * u, 0.0, 0.0, p / (g - 1) + 0.5 * rho * u * u])
    q = np.broadcast_to(qinf[:, None, None, None], (5, 16, 12, 10)).copy()
    m = cases.cartesian_metrics((16, 12, 10))
    topo = BlockTopology(0, {f: None for f in Face}, {f: BCType.FARFIELD for f in Face})
    cfg = Config(gas=GasModel(), scheme=SchemeChoice("muscl2", "roe", 3, "minmod"),
                 bc_params={f: {"q_inf": qinf} for f in Face}, topology=(topo,), ghost=2, nan_check=False)
    state = State(blocks=(Block(q=q, xyz=np.zeros((3, 16, 12, 10))),), metrics=(m,))
    for _ in range(3):
        state = step(state, cfg, 1e-6)
    assert state.step == 3 and state.t == pytest.approx(3e-6)
    np.testing.assert_allclose(state.blocks[0].q[:, 4:-4, 4:-4, 4:-4], q[:, 4:-4, 4:-4, 4:-4], rtol=1e-5, atol=1e-5)


def test_determinism_bitwise():
    print("case: determinism", flush=True)
    state, cfg = cases.case_c1(n=20, seed=2)
    outs = []
    for _ in range(2):
        st = Stepper(to_dev(state), cfg)
        st.advance(0.05, 4)
        torch.cuda.synchronize()
        outs.append(st.qs[0].cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.fixture(scope="module")
def c2_full():
    from famrdp_jax.core.types import Block, State
    from famrdp_jax.mesh.metric import compute_metrics
    q, xyz_ext, cfg = cases.case_c2(256)
    m = compute_metrics(torch.from_numpy(xyz_ext).cuda(), ghost=2)
    sl = (slice(2, -2),) * 3
    from famrdp_jax.core.types import Metrics
    mc = Metrics(jac=m.jac[sl].contiguous(), kxyz=m.kxyz[(slice(None), slice(None)) + sl].contiguous(),
                 vol=m.vol[sl].contiguous())
    from famrdp_jax.solver.cfl import cfl_dt
    dt, nbad = cfl_dt(torch.from_numpy(q).cuda(), mc, cfg)
    assert nbad == 0
    state = State(blocks=(Block(q=torch.from_numpy(q).cuda(), xyz=torch.zeros(1, dtype=torch.float64).cuda()),),
                  metrics=(mc,))
    return state, cfg, dt


def test_c2_full_size_center_subbox_vs_oracle(c2_full):
    """256^3: after one GPU step, a 26^3 sub-box around the centre matches
    the oracle's three stages on that sub-box (cells >= 7 from its faces are
    outside the influence of the crop: 3 stages x stencil radius 2)."""
    print("case: C2 256^3 centre sub-box, 1 step", flush=True)
    state, cfg, dt = c2_full
    st = Stepper(state, cfg)
    c0 = 117
    sl = (slice(None),) + (slice(c0, c0 + 26),) * 3
    q_before = st.qs[0][sl].cpu().numpy()
    msl = (slice(None), slice(None)) + (slice(c0, c0 + 26),) * 3
    k_sub = state.metrics[0].kxyz[msl].cpu().numpy()
    v_sub = state.metrics[0].vol[(slice(c0, c0 + 26),) * 3].cpu().numpy()
    st.advance(dt)
    torch.cuda.synchronize()
    got = st.qs[0][sl].cpu().numpy()
    from famrdp_jax.core.types import BlockTopology
    iso = BlockTopology(0, {}, {})
    ref = OB.step([q_before], [(k_sub, v_sub)], [iso], {}, cases.gas_dict(cfg.gas), dt)[0]
    inner = (slice(None),) + (slice(7, -7),) * 3
    err = rel_linf(got[inner], ref[inner])
    print(f"C2-256 sub-box rel Linf {err:.3e}", flush=True)
    assert err <= STEP_TOL


def test_c2_full_size_modes_agree_and_stay_finite(c2_full):
    print("case: C2 256^3 plain vs batched vs graph, 3 steps", flush=True)
    state, cfg, dt = c2_full
    outs = []
    for kw in (dict(), dict(batched=True), dict(graph=True)):
        st = Stepper(state, cfg)
        st.advance(dt, 3, **kw)
        torch.cuda.synchronize()
        outs.append(st.qs[0].clone())
    assert torch.isfinite(outs[0]).all()
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
