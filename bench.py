#!/usr/bin/env python3
"""Benchmark of the fused FP64 RK4 step on one MI355X.

Workload: (phi, Pi) wave at 256^3, 8th-order stencils (n=4), Kreiss-Oliger
order 4 with sigma=0.2, 16-cell quadratic absorbing layer (max rate 2.0) with
Dirichlet-0 ghosts on a non-periodic unit cube, dt = dx/4; initial data are 8
seeded Gaussians. A "step" is one RK4 step of the whole box (four fused stage
kernels), with the state resident in HBM; steps are replayed from a captured
hipGraph.

Prints one JSON line:
  value      cell-updates/s = 256^3 * steps / wall time between synchronisations
  roofline   algorithmic HBM bytes per stage launch (128 * ncomp B per cell per
             RK4 step, i.e. 32 * ncomp B per cell per launch on average) divided
             by the mean stage-kernel duration from HIP events recorded on the
             solver's stream around the timed steps, against 8 TB/s
  cpu_baseline  the CPU oracle (a restatement of the reference arithmetic,
             OpenMP over z planes) on the same 256^3 workload, a bounded number
             of steps on this host's cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
KERNEL_FIELDS = ("ghost_mode", "tile_y", "zchunk", "xcd_remap", "zalt", "halo_ring", "stage_pairs")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(cfg: dict):
    """HBM bytes per stage launch of stage-kernel configuration `cfg`, measured with rocprofv3
    counters (committed summary), or None if that configuration was not profiled."""
    p = ROOT / "profiles" / "fused_stage_wave256_traffic.json"
    if not p.exists():
        return None
    # single-stage configurations are keyed by their first six fields
    fields = KERNEL_FIELDS if cfg.get("stage_pairs") else KERNEL_FIELDS[:6]
    key = ",".join(str(cfg[k]) for k in fields)
    try:
        return json.loads(p.read_text())["configs"].get(key, {}).get("hbm_bytes_per_launch")
    except (ValueError, OSError, KeyError):
        return None


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, state, budget_s: float):
    """The C oracle on this host: RK4 steps of the same workload with all host threads for about
    `budget_s` seconds, then one step with one thread."""
    from oracle import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)

    def run(nthreads, budget, max_steps):
        orc.set_threads(nthreads)
        o = orc.OracleSolver(case)
        o.set_state(state)
        steps = 0
        t0 = time.perf_counter()
        while True:
            o.rk4_steps(1)
            steps += 1
            el = time.perf_counter() - t0
            if steps >= max_steps or el >= budget or el / steps * (steps + 1) > 2.0 * budget:
                break
        o.close()
        return steps, el

    steps, el = run(threads, budget_s, 1 << 30)
    steps1, el1 = run(1, 0.0, 1)
    orc.set_threads(threads)
    return {
        "value": case.cells * steps / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} RK4 step(s) of the same 256^3 workload ({case.cells} cells) by the C oracle, "
                  f"OpenMP {threads} threads, {el:.1f} s",
        "cpu_model": _cpu_model(),
        "single_thread_value": case.cells * steps1 / el1,
        "single_thread_sample": f"{steps1} RK4 step of the same workload on 1 thread, {el1:.1f} s",
    }


def ensemble_leg(count: int, steps: int, tune: bool):
    """Batched vs plain on the ensemble workload: `count` periodic (phi, Pi) instances (side in
    {32,48,64,96,128}, n in {2,3,4}, KO order n), `steps` RK4 steps each.
      batched: one Batch, graph-replayed steps, one persistent launch per stage per stencil group,
               each group with the configuration the library's autotuner chose for it;
      plain:   the instances one after another, each with the single-box solver and eager per-step
               launches, summed device time."""
    import torch

    from treesap_amd import Batch, Solver
    from treesap_amd.workloads import ensemble_params, gaussian_state_device

    params = ensemble_params(count=count)
    cells = sum(c.cells for c, _ in params)
    b = Batch([c for c, _ in params])
    for i, (case, centre) in enumerate(params):
        b.set_state(i, gaussian_state_device(case, centre))
    torch.cuda.synchronize()
    # the library's batch autotuner (tsp_batch_autotune): every candidate configuration timed on
    # this batch, the fastest kept; the state is restored
    best = (None, b.autotune(6) if tune else b.kernel)
    # every stencil-width group has its own configuration
    groups = sorted({(c.stencil_n, c.ko_order) for c, _ in params})
    group_cfgs = {f"n{sn}_ko{ko}": b.group_kernel(sn, ko) for sn, ko in groups}
    b.graph_rk4_steps(2)  # capture the graph of the chosen configuration (an even count keeps the parity)
    for i, (case, centre) in enumerate(params):
        b.set_state(i, gaussian_state_device(case, centre))
    batched_ms = b.time_rk4(steps, use_graph=True)
    torch.cuda.synchronize()

    # plain path: the same instances one after another; every instance must end
    # bitwise equal to its batched solve
    plain_ms = 0.0
    mismatched = 0
    for i, (case, centre) in enumerate(params):
        s = Solver(case)
        s.set_state(gaussian_state_device(case, centre))
        plain_ms += s.time_rk4(steps, use_graph=False)
        got = s.get_state(out=torch.empty(case.shape, dtype=torch.float64, device="cuda"))
        want = b.get_state(i, out=torch.empty(case.shape, dtype=torch.float64, device="cuda"))
        if not bool(torch.isfinite(want).all()) or not torch.equal(got, want):
            mismatched += 1
        s.close()
    b.close()
    if mismatched:
        raise RuntimeError(f"ensemble: {mismatched} instance(s) differ between the batched and plain solves")
    batched = cells * steps / (batched_ms / 1e3)
    plain = cells * steps / (plain_ms / 1e3)
    return {
        "workload": f"ensemble: {count} periodic (phi,Pi) instances, sides {{32,48,64,96,128}}, n in {{2,3,4}}, "
                    f"KO order n, sigma U[0.05,0.3], {steps} RK4 steps each",
        "instances": count,
        "cells": cells,
        "steps": steps,
        "batched_cell_updates_per_s": batched,
        "plain_cell_updates_per_s": plain,
        "batched_vs_plain": batched / plain,
        "batched_s": batched_ms / 1e3,
        "plain_s": plain_ms / 1e3,
        "batch_kernel": best[1],
        "batch_kernel_per_group": group_cfgs,
        "batch_kernel_chosen_by": "tsp_batch_autotune" if tune else "default",
        "batched_equals_plain": "bitwise, every instance",
    }


def tensor_leg(n: int, steps: int, tune: bool = True):
    """Linearised ADM (h_ij, K_ij) workload: 12 comps at n^3, n=3 stencils with mixed second
    derivatives, KO order 3, 12-cell ABL; graph-replayed steps."""
    import torch

    from treesap_amd import Solver
    from treesap_amd.workloads import fourier_state_device

    case = _tensor_case(n)
    s = Solver(case)
    s.set_state(fourier_state_device(case, "tensor192"))
    cfg = s.autotune(2) if tune else s.kernel  # block shape and z-slab depth of the block order
    s.graph_rk4_steps(2)
    ms = s.time_rk4(steps, use_graph=True)
    stage_ms, fill_ms = s.profile_rk4(2)
    st = s.get_state(out=torch.empty(case.shape, dtype=torch.float64, device="cuda"))
    if not bool(torch.isfinite(st).all()):
        raise RuntimeError("tensor leg: non-finite state after the timed steps")
    s.close()
    cu = case.cells * steps / (ms / 1e3)
    return {
        "workload": f"tensor{n}: linearised ADM (h_ij, K_ij) 12 comps {n}^3, n=3, KO order 3 sigma=0.1, "
                    f"12-cell ABL, {steps} RK4 steps timed",
        "cell_updates_per_s": cu,
        "algorithmic_GBs": cu * 128 * 12 / 1e9,
        "roofline_frac": cu * 128 * 12 / 8e12,
        "stage_ms": [x / 2 for x in stage_ms],
        "kernel": {"block": {4: "8x8x4", 16: "16x8x4"}.get(cfg["tile_y"], "16x8x2"), "slab_z_tiles": cfg["zchunk"],
                   "chosen_by": "tsp_autotune" if tune else "default"},
        "ghost_fill_ms_per_step": fill_ms / 2,
    }


def generic_leg(n: int, steps: int):
    """The stage kernel compiled from a user functor (include/treesap/device/generic.hpp), timed by the
    hipcc-built program tests/cpp/bin/generic_timing (a child process) on the wave256 arithmetic written
    as a functor the library does not carry: (phi, Pi) with n=4 stencils, KO order 4 sigma=0.2 and a
    16-cell quadratic ABL."""
    import subprocess

    prog = ROOT / "tests" / "cpp" / "bin" / "generic_timing"
    if not prog.exists():
        raise RuntimeError(f"{prog} is missing; build it with tests/cpp/build.py")
    res = subprocess.run([str(prog), str(n), str(steps)], capture_output=True, text=True, timeout=300)
    if res.returncode != 0:
        raise RuntimeError(f"generic_timing failed: {res.stderr.strip()}")
    out = json.loads(res.stdout.strip().splitlines()[-1])
    if not out["finite"]:
        raise RuntimeError("generic leg: non-finite state after the timed steps")
    out["workload"] = (f"wave{n} arithmetic as a user functor (KODecorator<ScalarWave<4,4>,4> under ABLPDEModel), "
                       "stage kernel compiled from the functor, graph-replayed RK4 steps")
    return out


def _tensor_case(n: int):
    from treesap_amd.workloads import Case

    return Case(system="tensor", n=(n, n, n), stencil_n=3, ghost=3, ko_order=3, ko_strength=0.1,
                periodic=(0, 0, 0), bc="dirichlet", abl_thickness=12, abl_max=2.0, abl_profile="quadratic",
                dt=0.25 * (1.0 / n))


def adjoint_leg(n: int, batch: int, nsteps: int, K: int):
    """Differentiable solve: (phi, Pi) n^3, n=2 stencils, 12-cell ABL, `batch` sources x `nsteps` RK4 steps,
    64 receivers; loss and gradient through the torch.autograd.Function (forward with receiver trace and
    checkpoints every K steps, backward through the transposed RK4 stage kernels), against the plain
    forward solve of the same sources (graph-replayed steps)."""
    import numpy as np
    import torch

    from treesap_amd import Solver
    from treesap_amd.adjoint import ReceiverProblem, receiver_loss
    from treesap_amd.workloads import adjoint128

    case, phi0, rec, target, nsteps = adjoint128(n=n, batch=batch, nsteps=nsteps)
    x = torch.tensor(phi0, device="cuda", requires_grad=True)
    tgt = torch.tensor(target, device="cuda")
    # plain forward of the same sources
    s = Solver(case)
    st = torch.zeros(case.shape, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in range(batch):
        st[0] = x.detach()[b]
        s.set_state(st)
        s.graph_rk4_steps(nsteps)
    s.sync()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    s.close()
    # the same forward solves as one batch (graph-replayed), for the stricter ratio
    from treesap_amd import Batch

    fb = Batch([case] * batch)
    for b in range(batch):
        st[0] = x.detach()[b]
        fb.set_state(b, st)
    fb.graph_prepare()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fb.graph_rk4_steps(nsteps)
    fb.sync()
    t_batched_fwd = time.perf_counter() - t0
    fb.close()
    prob = ReceiverProblem(case, rec, nsteps, batch=batch, checkpoint_every=K)
    out = {}
    for recompute in (False, True):
        x.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = receiver_loss(x, tgt, prob, recompute)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out["recompute" if recompute else "stored_trace"] = {
            "forward_s": t1 - t0, "backward_s": t2 - t1, "adjoint_over_forward": (t2 - t0) / t_plain,
            "adjoint_over_batched_forward": (t2 - t0) / t_batched_fwd}
        gnorm = float(x.grad.norm())
        if not (np.isfinite(float(loss.detach())) and bool(torch.isfinite(x.grad).all())):
            raise RuntimeError("adjoint leg: non-finite loss or gradient")
    prob.close()
    return {
        "workload": f"adjoint{n}: (phi,Pi) {n}^3, n=2, no KO, 12-cell ABL, {batch} sources x {nsteps} RK4 steps, "
                    f"64 receivers, checkpoints every {K}",
        "plain_forward_s": t_plain,
        "batched_forward_s": t_batched_fwd,
        "note": "adjoint = forward with receiver trace and checkpoints + backward, all sources in one batch; "
                "adjoint_over_forward divides by the plain single-box forward of the sources one after another, "
                "adjoint_over_batched_forward by the same forward solves as one graph-replayed batch",
        "loss": float(loss.detach()),
        "grad_norm": gnorm,
        **out,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="cells per side (default: the 256^3 workload)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU oracle stepping")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--kernel", default=None,
                    help="fixed stage-kernel config ghost_mode,tile_y,zchunk,xcd_remap[,zalt[,halo_ring]] (skips the autotuner)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-vs-plain ensemble leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-vs-forward leg")
    ap.add_argument("--no-tensor", action="store_true", help="skip the 12-component tensor leg")
    ap.add_argument("--no-generic", action="store_true", help="skip the user-functor (generic kernel) leg")
    ap.add_argument("--ensemble-count", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB (counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (this build has no "
              "multi-GPU path: each solve runs on a single MI355X)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from treesap_amd import Solver
    from treesap_amd.workloads import wave256

    case, state, _ = wave256(n=args.n)
    cells = case.cells
    log(f"[bench] workload (phi,Pi) {args.n}^3 n=4 KO4 sigma=0.2 ABL16, {cells} cells")

    s = Solver(case)
    s.set_state(state)
    if args.kernel:
        vals = [int(v) for v in args.kernel.split(",")]
        s.kernel = dict(zip(KERNEL_FIELDS, vals))
    elif not args.no_tune:
        t0 = time.perf_counter()
        cfg = s.autotune(3)
        log(f"[bench] autotune chose {cfg} in {time.perf_counter() - t0:.1f} s")
    cfg = s.kernel

    s.graph_rk4_steps(args.warmup)
    # graphs of both step parities exist before the timed region (an odd warmup
    # count would otherwise capture one inside it)
    s.graph_prepare()
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = s.time_rk4(args.steps, use_graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = s.get_state()
    if not np.isfinite(out).all():
        log("[bench] non-finite state after the timed steps")
        return 1

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy_mib} MiB x3 done")

    # The same steps with the state handed over in host memory: upload, step, download
    # (the rate a caller that keeps its fields on the host would see; not the headline value).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.set_state(state)
    s.graph_rk4_steps(args.steps)
    host_out = s.get_state()
    pcie_wall = time.perf_counter() - t0
    pcie = {
        "cell_updates_per_s": cells * args.steps / pcie_wall,
        "wall_s": pcie_wall,
        "note": "host numpy state uploaded, steps replayed, state downloaded to host, all inside the timed region",
    }
    del host_out

    # per-stage kernel durations (eager launches, events around each launch)
    stage_ms, fill_ms = s.profile_rk4(10)
    bytes_per_cell = s.rk4_bytes_per_cell()  # 128 * ncomp
    s.close()

    value = cells * args.steps / wall
    # four stage launches per step, or two with stage pairs (the algorithmic bytes of a pair
    # launch are those of its two stages)
    pairs = bool(cfg.get("stage_pairs"))
    launches = 2 if pairs else 4
    mean_launch_s = dev_ms / 1e3 / (launches * args.steps)
    bytes_per_launch = bytes_per_cell * cells / launches
    achieved = bytes_per_launch / mean_launch_s / 1e9
    # per launch kind: stages 0..3, or pairs (0,1) and (2,3) (profile_rk4 reports a pair in
    # the slot of its first stage)
    kinds = ((64, stage_ms[0]), (64, stage_ms[2])) if pairs else tuple(zip((24, 40, 40, 24), stage_ms))
    stage_gbs = [(b * case.ncomp * cells) / (ms / 1e3 / 10) / 1e9 for b, ms in kinds]
    log(f"[bench] {args.steps} steps: wall {wall * 1e3:.1f} ms, device {dev_ms:.1f} ms, "
        f"{value:.3e} cell-updates/s, per-stage GB/s {[round(x) for x in stage_gbs]}")

    ens = None
    if not args.no_ensemble:
        log(f"[bench] ensemble leg: {args.ensemble_count} instances x {args.ensemble_steps} steps ...")
        t0 = time.perf_counter()
        ens = ensemble_leg(args.ensemble_count, args.ensemble_steps, tune=not args.no_tune)
        log(f"[bench] ensemble: batched {ens['batched_cell_updates_per_s']:.3e} cu/s, plain "
            f"{ens['plain_cell_updates_per_s']:.3e} cu/s, ratio {ens['batched_vs_plain']:.2f} "
            f"({time.perf_counter() - t0:.1f} s)")

    ten = None
    if not args.no_tensor:
        log("[bench] tensor leg: 192^3 x 12 comps ...")
        t0 = time.perf_counter()
        ten = tensor_leg(192, 20, tune=not args.no_tune)
        log(f"[bench] tensor: {ten['cell_updates_per_s']:.3e} cu/s, {ten['algorithmic_GBs']:.0f} GB/s algorithmic "
            f"({100 * ten['roofline_frac']:.1f}%), stage ms {[round(x, 2) for x in ten['stage_ms']]} "
            f"({time.perf_counter() - t0:.1f} s)")

    gen = None
    if not args.no_generic:
        log("[bench] generic leg: the wave arithmetic as a user functor ...")
        t0 = time.perf_counter()
        gen = generic_leg(args.n, 20)
        gen["vs_registered_kernels"] = gen["cell_updates_per_s"] / value
        log(f"[bench] generic: {gen['cell_updates_per_s']:.3e} cu/s ({gen['vs_registered_kernels']:.2f} of the "
            f"registered kernels) ({time.perf_counter() - t0:.1f} s)")

    adj = None
    if not args.no_adjoint:
        log("[bench] adjoint leg: 128^3 x 16 sources x 500 steps ...")
        t0 = time.perf_counter()
        adj = adjoint_leg(128, 16, 500, 50)
        log(f"[bench] adjoint: plain forward {adj['plain_forward_s']:.3f} s, batched forward "
            f"{adj['batched_forward_s']:.3f} s; stored trace {adj['stored_trace']['adjoint_over_forward']:.2f}x "
            f"({adj['stored_trace']['adjoint_over_batched_forward']:.2f}x of batched); recompute "
            f"{adj['recompute']['adjoint_over_forward']:.2f}x ({adj['recompute']['adjoint_over_batched_forward']:.2f}x "
            f"of batched) ({time.perf_counter() - t0:.1f} s)")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU oracle baseline ...")
        cpu = cpu_baseline(case, state, args.cpu_budget)
        log(f"[bench] CPU: {cpu['value']:.3e} cell-updates/s ({cpu['sample']})")

    line = {
        "metric": "RK4 cell-updates/s & % HBM roofline (FP64); batched vs plain; adjoint/fwd time",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"wave256: (phi,Pi) wave {args.n}^3 FP64, 8th-order stencils (n=4), KO order 4 sigma=0.2, "
                        "16-cell quadratic ABL (max 2.0), Dirichlet 0, dt=dx/4, 8 seeded Gaussians",
            "cells": cells,
            "kernel": cfg,
            "graph": True,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(cfg),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "launches_per_step": launches,
            "mean_stage_launch_us": mean_launch_s * 1e6,
            "stage_gbs_eager": stage_gbs,
        },
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "batched": ens,
        "adjoint": adj,
        "tensor": ten,
        "generic": gen,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
