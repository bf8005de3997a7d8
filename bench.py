"""Benchmark: FP64 cell-updates/s of famrdp_jax's TVD-RK3 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]

Headline workload (``value``): the plain single-block step on the 256^3
curvilinear Taylor-Green case (260^3 padded, periodic, M0.1, Re 1600, CFL
0.5), state and metrics resident in HBM, K steps replayed from a captured
hipGraph, bracketed by device synchronisation.  A cell-update is one
interior cell advanced one full step (halo + BC + 3 RK stages).

Also reported on the same JSON line:
  roofline      live HIP-event timing of the RK3 stage kernels (eager launches
                on the launch stream) against 560 B of compulsory HBM traffic
                per padded cell per step (160 B stage 1, 200 B stages 2-3)
  cpu_baseline  the CPU oracle (numpy restatement of famrdp_jax's step, 1
                core) on a bounded sample of the same workload
  batched       the 64-block uneven multi-block case: one launch per kernel
                across all blocks (hipGraph) vs per-block plain launches
  adjoint       forward + adjoint over forward cost on the wall-bounded case
                and on the 32-case ensemble
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "famrdp-jax_amd"))

METRIC = "FP64 cell-updates/s per RK3 step (plain vs batched) + adjoint/forward cost, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BYTES_PER_PADDED_CELL_STEP = 560.0   # 160 (stage 1) + 200 + 200
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_kernel_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _sync():
    import torch
    torch.cuda.synchronize()


def setup_c2(n):
    from famrdp_jax import cases
    return setup_curvilinear(*cases.case_c2(n))


def setup_curvilinear(q, xyz_ext, cfg):
    """Metrics on the GPU from coordinates with 2 extra layers, cropped."""
    import torch
    from famrdp_jax.core.types import Block, Metrics, State
    from famrdp_jax.mesh.metric import compute_metrics
    from famrdp_jax.solver.cfl import cfl_dt
    m = compute_metrics(torch.from_numpy(xyz_ext).cuda(), ghost=2)
    del xyz_ext
    sl = (slice(2, -2),) * 3
    mc = Metrics(jac=m.jac[sl].contiguous(), kxyz=m.kxyz[(slice(None), slice(None)) + sl].contiguous(),
                 vol=m.vol[sl].contiguous())
    del m
    qd = torch.from_numpy(q).cuda()
    dt, nbad = cfl_dt(qd, mc, cfg)
    if nbad:
        raise FloatingPointError("initial state has invalid cells")
    state = State(blocks=(Block(q=qd, xyz=torch.zeros(1, dtype=torch.float64, device="cuda")),), metrics=(mc,))
    return state, cfg, dt


def bench_c2(n, steps, warmup, timing_steps=3, tile=-1):
    from famrdp_jax.solver.stepper import Stepper
    state, cfg, dt = setup_c2(n)
    st = Stepper(state, cfg)
    npad = (n + 4) ** 3
    if tile >= 0:
        st.plan.set_tile(tile)
        v, dims = st.plan.tile()
        log(f"C2 {n}^3: dt={dt:.4e}; stage-kernel variant {v} (tile {dims}) as requested")
    else:
        v, dims, ms = st.autotune(dt)
        log(f"C2 {n}^3: dt={dt:.4e}; autotuned tile {dims} (variant {v}); stage 1+2 ms per variant "
            + " ".join(f"{x:.3f}" for x in ms))
    st.advance(dt, warmup, graph=True)
    _sync()
    t0 = time.perf_counter()
    st.advance(dt, steps, graph=True)
    _sync()
    wall = time.perf_counter() - t0
    log(f"C2 timed {steps} steps: {wall * 1e3 / steps:.3f} ms/step")
    # live per-kernel timing (eager launches, events on the launch stream)
    st.plan.timing_start()
    st.advance(dt, timing_steps)
    stage_ms, nstage = st.plan.timing_stop()
    _sync()
    import torch
    finite = bool(torch.isfinite(st.qs[0]).all().item())
    avg_stage_ms = stage_ms / max(nstage, 1)
    bytes_per_stage = npad * BYTES_PER_PADDED_CELL_STEP / 3.0
    achieved = bytes_per_stage / (avg_stage_ms * 1e-3) / 1e9
    return dict(n=n, npad=npad, dt=dt, steps=steps, wall=wall, ms_per_step=wall * 1e3 / steps, tile=dims, variant=v,
                value=n ** 3 * steps / wall, avg_stage_ms=avg_stage_ms, stage_bytes=bytes_per_stage,
                achieved_gbs=achieved, finite=finite)


def cpu_baseline(n=112, target_s=10.0):
    """The reference's step on the host CPU: the oracle's C + OpenMP
    restatement on every host thread (oracle/c_step.py, pinned against the
    numpy oracle) on the same curvilinear TGV workload at n^3 interior; one
    warm-up step, then as many steps as fill about ``target_s`` seconds.
    Without the built C oracle, the numpy oracle on one core."""
    import numpy as np
    from famrdp_jax import cases
    from oracle import blocks as OB
    from oracle import c_step
    from oracle import metric as OM
    q, xyz_ext, cfg = cases.case_c2(n)
    jac, k, vol = OM.compute_metrics(xyz_ext)
    m = cases.crop_metrics(jac, k, vol)
    gas = cases.gas_dict(cfg.gas)
    dt = OM.cfl_dt(q, m.kxyz, m.vol, gas, 0.5)
    if c_step.available():
        # the job's CPU share: OMP_NUM_THREADS where set, else at most 16 threads
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or min(16, os.cpu_count() or 1)
        one = lambda x: c_step.step(x, m.kxyz, m.vol, cfg.topology[0], gas, dt, threads=threads)
        what, cores = f"oracle C + OpenMP step ({threads} threads)", threads
    else:
        mets, topo = [(m.kxyz, m.vol)], list(cfg.topology)
        one = lambda x: OB.step([x], mets, topo, cfg.bc_params, gas, dt)[0]
        what, cores = "oracle numpy step (C oracle not built)", 1
    t0 = time.perf_counter()
    q = one(q)                                   # warm-up, also sizes the sample
    steps = max(3, min(50, int(target_s / max(time.perf_counter() - t0, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(steps):
        q = one(q)
    wall = time.perf_counter() - t0
    assert np.isfinite(q).all()
    return dict(value=n ** 3 * steps / wall, unit="cell-updates/s", cores=cores, kind="port",
                sample=f"{what} on the same curvilinear TGV workload at {n}^3 interior ({n + 4}^3 padded), "
                       f"{steps} steps after 1 warm-up, {wall:.1f} s",
                host_cpus=os.cpu_count(), omp_threads=os.environ.get("OMP_NUM_THREADS"))


def bench_c3(steps=10, warmup=2):
    import torch
    from famrdp_jax import cases
    from famrdp_jax.solver.cfl import cfl_dt
    from famrdp_jax.solver.stepper import Stepper, state_to_device
    state, cfg = cases.case_c3()
    dev = state_to_device(state)
    dt = min(cfl_dt(b.q, m, cfg)[0] for b, m in zip(dev.blocks, dev.metrics))
    ncell = sum((b.q.shape[1] - 4) * (b.q.shape[2] - 4) * (b.q.shape[3] - 4) for b in dev.blocks)
    res = dict(blocks=len(dev.blocks), interior_cells=int(ncell), dt=dt)
    for name, kw in (("batched_graph", dict(batched=True, graph=True)),
                     ("per_block_graph", dict(batched=False, graph=True)),
                     ("per_block_plain", dict(batched=False, graph=False))):
        st = Stepper(dev, cfg)
        v, dims, _ = st.autotune(dt, batched=kw["batched"])
        st.advance(dt, warmup, **kw)
        _sync()
        t0 = time.perf_counter()
        st.advance(dt, steps, **kw)
        _sync()
        wall = time.perf_counter() - t0
        res[name] = dict(ms_per_step=wall * 1e3 / steps, cell_updates_per_s=ncell * steps / wall, tile=dims)
        log(f"C3 {name}: {wall * 1e3 / steps:.3f} ms/step")
        del st
        torch.cuda.empty_cache()
    res["speedup_batched_vs_per_block_plain"] = (res["per_block_plain"]["ms_per_step"]
                                                 / res["batched_graph"]["ms_per_step"])
    return res


def _time_adjoint(adj, dt, nsteps, kind, k):
    """forward alone vs forward + adjoint, checkpointing every k steps and with
    every step's states kept in memory (k = 0)."""
    adj.forward(dt, 2)
    _sync()
    t0 = time.perf_counter()
    adj.forward(dt, nsteps)
    _sync()
    fwd = time.perf_counter() - t0
    tm, tm0 = {}, {}
    adj.gradient(dt, nsteps, kind, checkpoint_every=k, timing=tm)
    adj.gradient(dt, nsteps, kind, checkpoint_every=0, timing=tm0)
    v, dims = adj.plan.tile()
    return dict(steps=nsteps, checkpoint_every=k, forward_ms=fwd * 1e3, gradient_ms=tm["total_s"] * 1e3,
                ratio=tm["total_s"] / fwd, gradient_all_states_in_memory_ms=tm0["total_s"] * 1e3,
                ratio_all_states_in_memory=tm0["total_s"] / fwd, forward_variant=v, forward_tile=list(dims))


def bench_adjoint():
    """forward + adjoint cost over forward cost (eager launches both):
    C4 (96^3 curvilinear wall/farfield, 50 steps, wall-pressure J, checkpoint
    every 10) and C5 (32 periodic cases in one batched plan with their own
    mu0 and dt, 30 steps, kinetic-energy J)."""
    import torch
    from famrdp_jax import cases
    from famrdp_jax.adjoint import OBJ_KINETIC_ENERGY, OBJ_WALL_PRESSURE, Adjoint
    from famrdp_jax.solver.cfl import cfl_dt
    from famrdp_jax.solver.stepper import state_to_device
    res = {}
    state, cfg, dt = setup_curvilinear(*cases.case_c4(96))
    adj = Adjoint(state, cfg, autotune=True)
    res["c4"] = _time_adjoint(adj, dt, 50, OBJ_WALL_PRESSURE, 10)
    log(f"C4 adjoint: {res['c4']}")
    del adj, state
    torch.cuda.empty_cache()
    pairs = cases.case_c5(32)
    state, cfg, mu0 = cases.ensemble(pairs)
    dev = state_to_device(state)
    dts = [cfl_dt(b.q, m, c)[0] for b, m, (_, c) in zip(dev.blocks, dev.metrics, pairs)]
    adj = Adjoint(dev, cfg, batched=True, autotune=True)
    adj.set_block_params(mu0, dts)
    res["c5"] = _time_adjoint(adj, 1.0, 30, OBJ_KINETIC_ENERGY, 10)
    res["c5"]["cases"] = len(pairs)
    log(f"C5 adjoint: {res['c5']}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the C4/C5 forward+adjoint measurement (its own JSON line; for profiling the reverse)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the C3 batched vs per-block measurement (its own JSON line; for profiling it)")
    ap.add_argument("--tile", type=int, default=-1, help="stage-kernel variant (default: autotune)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 2 GiB float64 buffer 5 times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the step would shard by "
              "blocks with one halo exchange per RK3 step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        a = torch.ones(2 ** 28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        _sync()
        log("copy probe: 5 copies of 2 GiB (read 2 GiB + write 2 GiB each)")
        return
    from famrdp_jax import _lib
    _lib.load()
    if args.adjoint_only:
        print(json.dumps(dict(adjoint=bench_adjoint())), flush=True)
        return
    if args.batched_only:
        print(json.dumps(dict(batched=bench_c3())), flush=True)
        return
    c2 = bench_c2(args.n, args.steps, args.warmup, tile=args.tile)
    # HBM bytes per stage launch of the chosen variant, measured with rocprofv3
    # counters for this library version and size (profiles/stage_kernel_traffic.json)
    # and the FP64 operations it executes (SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64, FMA counted twice)
    traffic = flops = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            t = json.load(open(TRAFFIC_FILE))
            if t.get("version") == _lib.lib().famrdp_version().decode() and t.get("n") == args.n:
                traffic = t.get("hbm_bytes_per_stage", {}).get(str(c2["variant"]))
                flops = t.get("fp64_flop_per_stage", {}).get(str(c2["variant"]))
        except Exception:
            traffic = flops = None
    roofline = dict(bound="hbm", achieved=round(c2["achieved_gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(c2["achieved_gbs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel="RK3 stage (" + _lib.lib().famrdp_version().decode() + ")",
                    bytes_per_launch=c2["stage_bytes"], avg_launch_ms=round(c2["avg_stage_ms"], 4))
    if flops:
        tf = flops / (c2["avg_stage_ms"] * 1e-3) / 1e12
        roofline.update(fp64_flop_per_launch=flops, fp64_tflops=round(tf, 2), fp64_peak_tflops=FP64_PEAK_TFLOPS,
                        fp64_frac=round(tf / FP64_PEAK_TFLOPS, 4))
    out = dict(metric=METRIC, value=round(c2["value"], 1), unit="cell-updates/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(c2["ms_per_step"], 4), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload=f"C2 plain step: single-block curvilinear {args.n}^3 TGV "
                                    f"({args.n + 4}^3 padded), periodic, M0.1, Re1600, CFL0.5, hipGraph replay",
                           global_batch=1, seq_len=None, parallelism="single GPU", interior_cells=args.n ** 3,
                           dt=c2["dt"], finite_after=c2["finite"], tile=c2["tile"]),
               roofline=roofline)
    if not args.no_batched:
        try:
            out["batched"] = bench_c3()
        except Exception as e:  # report, do not hide the headline
            out["batched"] = dict(error=repr(e))
    if not args.no_adjoint:
        try:
            out["adjoint"] = bench_adjoint()
        except Exception as e:
            out["adjoint"] = dict(error=repr(e))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
