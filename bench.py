#!/usr/bin/env python3
"""Benchmark: FP64 Multipole potential+force throughput on one MI355X.

Workload: Multipole halo, lmax=8, 50 log radial nodes on [0.1, 500] kpc, from
a triaxial Dehnen density (gamma=1, M=1e12, a=20 kpc, axes 1:0.8:0.6; 15
active harmonics after symmetry pruning), potential+force at N=10M positions
(Hernquist a=20 kpc, seed 0, 1% log-uniform in [0.01, 2000] kpc).  Inputs and
tables are resident in HBM before the timed region; one step = one launch
over all N particles.

Prints one JSON line (rank 0).  Extra keys: ``roofline`` (FP64 VALU bound,
algorithmic flops per launch / the kernel's average launch duration, timed by
one pair of HIP events around the K steps on the launch stream; HBM traffic
per launch from the committed rocprof summary) and ``cpu_baseline`` (the oracle's OpenMP C
restatement on this host's cores, same workload), and the other legs of the
metric (skipped with --no-extras): ``dense`` (the same points through the
dense 81-term lmax-8 FIRE table), ``analytic`` (NFW, HBM-bound),
``composite`` (Milky-Way-like Multipole + CylSpline + NFW + Plummer, evalDeriv
at N=1M in one fused launch), ``high_order`` (lmax-32 Multipole evalDeriv at
N=10M), ``batched`` (256-potential ragged ensemble, one launch vs one per
potential), ``grad`` (force-fit forward+backward) and ``leapfrog`` (HIP-graph
replay vs per-step API calls).  ``--only LEG`` runs one of them alone.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector peak
PEAK_HBM_GBS = 8000.0


def opcount_multipole(desc: dict, kinds: np.ndarray, order: int = 1) -> float:
    """Algorithmic FP64 flops (FMA = 2, each transcendental/sqrt/divide = 1)
    for one launch, summed over particles by the branch each one takes.

    Interior point, pot+force: 80 + 40 n + 12 S + 5 G, with n active (l,m)
    terms including (0,0), G active |m| groups and S = sum over groups of
    (l_last - |m|) recurrence steps; pot only: 60 + 16 n + 6 S; evalDeriv:
    350 + 90 n + 12 S + 5 G.  Extrapolated point: 40 (pot+force), 60
    (evalDeriv), 20 (pot).  The count is that of the one-step recurrence
    for every table; the kernels step by 2 in groups whose terms all have
    l - |m| even, and run S_executed steps (reported beside S).
    """
    prog = desc["program"]
    n = desc["n_lm"]
    G = len(prog.grp_m)
    S = int(sum(int(prog.step_l[prog.grp_step0[g + 1] - 1]) - int(prog.grp_m[g]) for g in range(G)))
    S_exec = int(np.sum(np.diff(prog.grp_step0) - 1))
    if order == 0:
        f_in, f_ex = 60 + 16 * n + 6 * S, 20
    elif order == 1:
        f_in, f_ex = 80 + 40 * n + 12 * S + 5 * G, 40
    else:
        f_in, f_ex = 350 + 90 * n + 12 * S + 5 * G, 60
    n_in = int(np.count_nonzero(kinds == 2))
    return float(f_in * n_in + f_ex * (len(kinds) - n_in)), dict(n=n, G=G, S=S, S_executed=S_exec,
                                                                   f_interior=f_in, f_extrap=f_ex,
                                                                   n_interior=n_in)


def classify(desc: dict, x: np.ndarray) -> np.ndarray:
    r = np.sqrt((x * x).sum(1))
    kinds = np.full(len(x), 2, np.int8)
    with np.errstate(divide="ignore"):
        lr = np.log(r)
    kinds[lr < desc["logr_min"]] = 1
    k = np.floor((lr - desc["logr_min"]) * desc["inv_dlogr"])
    kinds[(kinds == 2) & (k >= desc["n_intervals"])] = 3
    kinds[r < 1e-300] = 0
    return kinds


def profiled_value(workload_file: str, key: str):
    """A figure recorded in a committed rocprof summary as ``key = value``:
    HBM bytes per launch from the copy-calibrated FETCH_SIZE / WRITE_SIZE runs
    (traffic_bytes_per_launch), or FP64 flops executed per launch from the
    SQ_INSTS_VALU_*_F64 counters (fp64_flops_per_launch)."""
    path = os.path.join(ROOT, "profiles", workload_file)
    try:
        for ln in open(path):
            if ln.startswith(key):
                return int(float(ln.split("=", 1)[1]))
    except OSError:
        pass
    return None


def cpu_share() -> dict:
    """The host cores this process may use: its affinity mask, the cgroup CPU
    quota (cpu.max, in cores) when one is set, and OMP_NUM_THREADS when set;
    ``threads`` is the smallest of them (the CPU baseline's thread count)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    omp = int(omp) if omp and omp.isdigit() else None
    cands = [aff] + [math.ceil(quota)] * (quota is not None) + [omp] * (omp is not None)
    return {"threads": max(1, min(cands)), "affinity_cores": aff, "cgroup_quota_cores": quota,
            "omp_num_threads": omp, "nproc": os.cpu_count()}


def cpu_threads() -> int:
    return cpu_share()["threads"]


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_batched(dev, reps: int = 5):
    """Ragged ensemble of 256 Multipole potentials (seed 3: lmax 4..12, 30..80
    nodes, triaxial Dehnen tables, counts log-uniform in [1e3, 1e6]), pot+force:
    one persistent launch vs one plain launch per potential, both timed over
    `reps` passes with inputs resident."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    from agama_gpu_amd.workloads import ensemble_coefs, ensemble_spec, hernquist_points

    t_setup = time.perf_counter()
    spec = ensemble_spec(256, 3)
    with ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
        coefs = list(ex.map(ensemble_coefs, spec))
    pots = [MultipolePotentialGPU(c, device=dev) for c in coefs]
    counts = [s["count"] for s in spec]
    n = int(sum(counts))
    xyz = torch.as_tensor(hernquist_points(n, 3), device=dev)
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    force = torch.empty((n, 3), dtype=torch.float64, device=dev)
    t_tune = time.perf_counter()
    ens = MultipoleEnsembleGPU(pots, counts, tune=True)
    t_ready = time.perf_counter()
    stream = torch.cuda.current_stream(dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - w0) * 1e3 / reps, a.elapsed_time(b) / reps

    wall_b, gpu_b = timed(lambda: ens.eval_into(xyz, phi, force))
    ref = force.clone()
    wall_s, gpu_s = timed(lambda: ens.eval_separately(xyz, phi, force))
    agree = float((torch.max(torch.abs(force - ref)) / torch.max(torch.abs(ref))).item())
    # the per-potential launches captured in one HIP graph and replayed: the
    # same kernels without the host's launch gaps between them
    side = torch.cuda.Stream(dev)
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        ens.eval_separately(xyz, phi, force)
    stream.wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ens.eval_separately(xyz, phi, force)
    force.fill_(0.0)
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    agree_graph = float((torch.max(torch.abs(force - ref)) / torch.max(torch.abs(ref))).item())
    _, gpu_g = timed(g.replay)
    paths = {}
    for (pth, rows) in ens.choices:
        k = f"{['lane', 'lds'][pth]}/{rows}"
        paths[k] = paths.get(k, 0) + 1
    return {"n_potentials": len(pots), "n_particles": n, "work_items": int(len(ens.items)),
            "ms_batched": wall_b, "ms_batched_events": gpu_b,
            "ms_per_potential_launches": wall_s, "ms_per_potential_events": gpu_s,
            "speedup": wall_s / wall_b, "particles_per_s_batched": n / (wall_b * 1e-3),
            "ms_per_potential_graph": gpu_g, "speedup_vs_graph": gpu_g / gpu_b,
            "max_rel_diff_vs_per_potential": agree, "max_rel_diff_graph": agree_graph,
            "choices": paths,
            "setup_s": t_tune - t_setup, "tune_s": t_ready - t_tune}


def _event_times(fn, reps: int = 20, settle: float = 0.5):
    """HIP-event times of `reps` launches of fn, after `settle` seconds of
    untimed launches: a leg that follows the CPU-baseline leg (seconds with the
    GPU idle) otherwise starts at low clocks and reads up to 20 % slow."""
    import torch
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    while time.perf_counter() - t0 < settle:
        fn()
        torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in evs]


def bench_dense(dev, x_np, xyz, phi, force):
    """The same N=10M points through the dense lmax-8 table of the reference's
    FIRE fixture (600.dark.none_8.coef_mul_DR: all 81 harmonics, 25 nodes),
    pot+force: the table the FP64-VALU roofline target is stated for."""
    from agama_gpu_amd import MultipolePotentialGPU, read_coefs
    coefs = read_coefs(os.path.join(ROOT, "tests", "golden", "600.dark.none_8.coef_mul_DR"))
    pot = MultipolePotentialGPU(coefs, device=dev)
    kinds = classify(pot._data, x_np)
    flops, info = opcount_multipole(pot._data, kinds, order=1)
    ts = _event_times(lambda: pot.eval_into(xyz, phi=phi, force=force))
    ms = float(np.median(ts))
    ach = flops / (ms * 1e-3) / 1e12
    return {"table": "600.dark.none_8.coef_mul_DR (81 terms, 25 nodes)", "kernel_ms_median": ms,
            "particles_per_s": len(x_np) / (ms * 1e-3), "achieved_tflops": ach,
            "frac_fp64_peak": ach / PEAK_FP64_TFLOPS, "opcount": info}


def bench_analytic(dev, xyz, phi, force):
    """NFW (M=1e12, rs=20 kpc) pot+force at the same points: the HBM-bound
    member of the path (56 B per particle)."""
    from agama_gpu_amd import NFWPotentialGPU
    pot = NFWPotentialGPU(mass=1e12, scaleRadius=20.0, device=dev)
    n = xyz.shape[0]
    ts = _event_times(lambda: pot.eval_into(xyz, phi=phi, force=force))
    ms = float(np.median(ts))
    gbs = 56.0 * n / (ms * 1e-3) / 1e9
    return {"kernel_ms_median": ms, "particles_per_s": n / (ms * 1e-3), "achieved_gbs": gbs,
            "frac_hbm_peak": gbs / PEAK_HBM_GBS}


def bench_composite(dev, n: int = 1_000_000):
    """Milky-Way-like composite (Multipole halo lmax 8 + CylSpline disk and
    bar, 32 x 32 grid, mmax 6 + NFW + Plummer), evalDeriv (phi, force,
    Hessian) at N=1M Hernquist points in one fused launch."""
    import torch
    from agama_gpu_amd.workloads import hernquist_points, mw_composite
    comp = mw_composite(device=dev)
    xyz = torch.as_tensor(hernquist_points(n, 11), device=dev)
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    force = torch.empty((n, 3), dtype=torch.float64, device=dev)
    deriv = torch.empty((n, 6), dtype=torch.float64, device=dev)
    ts = _event_times(lambda: comp.eval_into(xyz, phi=phi, force=force, deriv=deriv))
    ms = float(np.median(ts))
    return {"components": "Multipole lmax 8 + CylSpline 32x32 mmax 6 + NFW + Plummer",
            "n_particles": n, "kernel_ms_median": ms, "particles_per_s": n / (ms * 1e-3),
            "launches": 1}


def bench_cylspline(dev, n: int = 1_000_000):
    """The CylSpline disk + bar table of the Milky-Way-like composite alone
    (Miyamoto-Nagai disk plus m = 2, 4, 6 bar, 32 x 32 asinh grid mirrored
    to 32 x 63, 7 harmonics) at N=1M Hernquist points (seed 11): pot+force and evalDeriv,
    with the algorithmic flop counts 150 + 150 n_harm and 250 + 220 n_harm
    per point."""
    import torch
    from agama_gpu_amd import CylSplinePotentialGPU
    from agama_gpu_amd.workloads import hernquist_points, mw_disk_bar_coefs
    pot = CylSplinePotentialGPU(mw_disk_bar_coefs(), device=dev)
    nh = int(pot._data["n_harm"])
    xyz = torch.as_tensor(hernquist_points(n, 11), device=dev)
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    force = torch.empty((n, 3), dtype=torch.float64, device=dev)
    deriv = torch.empty((n, 6), dtype=torch.float64, device=dev)
    out = {"n_particles": n, "n_harm": nh, "grid": [int(pot._data["lR_grid"].size),
                                                    int(pot._data["lz_grid"].size)]}
    for name, fn, fl in (("pot_force", lambda: pot.eval_into(xyz, phi=phi, force=force),
                          150 + 150 * nh),
                         ("evalderiv", lambda: pot.eval_into(xyz, phi=phi, force=force,
                                                             deriv=deriv), 250 + 220 * nh)):
        ms = float(np.median(_event_times(fn)))
        tf = fl * n / (ms * 1e-3) / 1e12
        out[name] = {"kernel_ms_median": ms, "particles_per_s": n / (ms * 1e-3),
                     "flops_per_point": fl, "achieved_tflops": tf,
                     "frac_fp64_peak": tf / PEAK_FP64_TFLOPS}
    return out


def bench_high_order(dev, n: int = 10_000_000):
    """Multipole lmax 32 (80 nodes, 1088 active harmonics), evalDeriv at
    N=10M points log-uniform in r over [0.05, 800] kpc."""
    import torch
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.workloads import high_order_coefs, loguniform_points
    pot = MultipolePotentialGPU(high_order_coefs(), device=dev)
    x_np = loguniform_points(n, 2, 0.05, 800.0)
    flops, info = opcount_multipole(pot._data, classify(pot._data, x_np), order=2)
    xyz = torch.as_tensor(x_np, device=dev)
    del x_np
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    force = torch.empty((n, 3), dtype=torch.float64, device=dev)
    deriv = torch.empty((n, 6), dtype=torch.float64, device=dev)
    ts = _event_times(lambda: pot.eval_into(xyz, phi=phi, force=force, deriv=deriv), reps=10)
    ms = float(np.median(ts))
    ach = flops / (ms * 1e-3) / 1e12
    return {"n_particles": n, "kernel_ms_median": ms, "particles_per_s": n / (ms * 1e-3),
            "achieved_tflops": ach, "frac_fp64_peak": ach / PEAK_FP64_TFLOPS, "opcount": info}


def bench_grad(dev, reps: int = 20, warmup: int = 10):
    """Force fit: Multipole lmax 8 (2 x 81 x 50 live table entries) shifted
    along a 200-knot spline centre, N=1M tracers; one step = loss (mean
    squared force residual, sum over particles and components / N, torch's
    fused squared-error sum) forward + backward to the tables, knots and
    positions."""
    import torch
    from agama_gpu_amd.workloads import fit_problem
    model, xyz, F_t = fit_problem(1_000_000, 200, device=dev)
    xyz = xyz.clone().requires_grad_(True)
    params = [model.multipole.phi_table, model.multipole.dphi_table, model.center.knots, xyz]
    n = xyz.shape[0]

    def step():
        for q in params:
            q.grad = None
        F = model(xyz, 0.73)[1]
        loss = torch.nn.functional.mse_loss(F, F_t, reduction="sum") / n
        loss.backward()
        return loss

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - w0) * 1e3 / reps
    with torch.no_grad():
        t0 = time.perf_counter()
        for _ in range(reps):
            model(xyz, 0.73)
        torch.cuda.synchronize()
        ms_fwd = (time.perf_counter() - t0) * 1e3 / reps
    return {"n_particles": int(xyz.shape[0]), "table_entries": int(2 * model.multipole.phi_table.numel()),
            "knots": int(model.center.knots.shape[0]), "ms_forward_backward": ms,
            "ms_forward": ms_fwd, "particles_per_s_grad": xyz.shape[0] / (ms * 1e-3)}


def bench_leapfrog(dev, n: int = 1_000_000, n_steps: int = 1000, eager_steps: int = 50):
    """Leapfrog of n test particles for n_steps through an Evolving halo of 8
    lmax-8 snapshots plus a spline-shifted lmax-4 satellite: HIP-graph replay
    (50 steps per graph, device-resident step counter) vs the same steps
    through the public API one at a time (time from the host), the latter on
    the first `eager_steps` steps."""
    import torch
    from agama_gpu_amd import LeapfrogGPU
    from agama_gpu_amd.workloads import leapfrog_problem
    pot, x, v = leapfrog_problem(n, device=dev)
    lf = LeapfrogGPU(pot, dt=1e-3, n_steps=n_steps)
    xt = torch.as_tensor(x, device=dev)
    vt = torch.as_tensor(v, device=dev)
    lf.run(xt, vt, n_steps=50, graph=True, steps_per_graph=50)     # capture + warm-up
    torch.cuda.synchronize()
    lf.set_state(xt, vt)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    lf.advance(n_steps, graph=True, steps_per_graph=50)
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - w0) * 1e3 / n_steps
    xg = lf.x.clone()
    lf.run_eager(xt, vt, 2)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    xe, _ = lf.run_eager(xt, vt, eager_steps)
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - w0) * 1e3 / eager_steps
    return {"n_particles": n, "n_steps": n_steps, "leaves_per_step": lf.max_active,
            "ms_per_step_graph": ms_graph, "ms_per_step_eager": ms_eager,
            "speedup": ms_eager / ms_graph, "particle_steps_per_s": n / (ms_graph * 1e-3),
            "final_state_finite": bool(torch.isfinite(xg).all().item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--settle", type=float, default=0.5,
                    help="seconds of untimed steps before the warm-up steps (GPU clock ramp)")
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-ensemble and gradient legs")
    ap.add_argument("--only", choices=["dense", "analytic", "composite", "cylspline", "high_order",
                                       "batched", "grad", "leapfrog"],
                    help="run one secondary leg alone and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 8}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): evaluation of "
              "independent particles shards by contiguous ranges with replicated tables; "
              "no multi-GPU figure is produced here.", file=sys.stderr)
        sys.exit(2)

    import torch
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.workloads import halo_lmax8_coefs, hernquist_points

    if args.only in ("composite", "cylspline", "high_order", "batched", "grad", "leapfrog"):
        dev = torch.device("cuda", 0)
        leg = {"composite": bench_composite, "cylspline": bench_cylspline,
               "high_order": bench_high_order, "batched": bench_batched, "grad": bench_grad,
               "leapfrog": bench_leapfrog}[args.only]
        print(json.dumps({args.only: leg(dev)}), flush=True)
        return

    N = args.n
    coefs = halo_lmax8_coefs()
    x_np = hernquist_points(N, 0)
    dev = torch.device("cuda", 0)
    pot = MultipolePotentialGPU(coefs, device=dev)
    desc = pot._data
    kinds = classify(desc, x_np)
    flops, opinfo = opcount_multipole(desc, kinds, order=1)

    xyz = torch.as_tensor(x_np, device=dev)
    phi = torch.empty(N, dtype=torch.float64, device=dev)
    force = torch.empty((N, 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    if args.only:
        res = (bench_dense(dev, x_np, xyz, phi, force) if args.only == "dense"
               else bench_analytic(dev, xyz, phi, force))
        print(json.dumps({args.only: res}), flush=True)
        return

    def step():
        pot.eval_into(xyz, phi=phi, force=force)

    # Clock settle: the GPU ramps its clocks over the first ~100 ms of load
    # (5 warm-up launches leave the kernel 25 % slower than steady state), so
    # the step runs untimed for `settle` seconds before the W warm-up steps.
    t_settle = time.perf_counter()
    n_settle = 0
    while time.perf_counter() - t_settle < args.settle:
        step()
        torch.cuda.synchronize()
        n_settle += 1
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    # The timed region: K back-to-back steps between one pair of HIP events on
    # the launch stream (the kernel's average launch duration) and the wall
    # clock (the step time, host launch included).
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    k_mean = e0.elapsed_time(e1) / args.steps
    # per-launch event pairs after the region, for the median (each pair also
    # times the gap before its launch)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    k_median = float(np.median([a.elapsed_time(b) for a, b in evs]))
    value = N / (ms_step * 1e-3)
    achieved = flops / (k_mean * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import OracleMultipole
        orc = OracleMultipole.from_coefs(coefs)
        th = cpu_threads()
        for _ in range(2):
            orc.pot_force(x_np, th)
        ts = []
        for _ in range(args.cpu_reps):
            c0 = time.perf_counter()
            orc.pot_force(x_np, th)
            ts.append(time.perf_counter() - c0)
        t_cpu = float(np.median(ts))
        share = cpu_share()
        cpu = {"value": N / t_cpu, "unit": "particles/s", "cores": th, "kind": "port",
               **{k: v for k, v in share.items() if k != "threads"},
               "sample": f"full N={N} workload, median of {args.cpu_reps} after 2 warmups; "
                         f"oracle/mp_oracle.c -O3 OpenMP, {th} threads (the smallest of the "
                         f"affinity mask, the cgroup quota and OMP_NUM_THREADS) on "
                         f"{cpu_model()}"}

    extras = {}
    if not args.no_extras:
        extras["dense"] = bench_dense(dev, x_np, xyz, phi, force)
        extras["analytic"] = bench_analytic(dev, xyz, phi, force)
        extras["composite"] = bench_composite(dev)
        extras["cylspline"] = bench_cylspline(dev)
        extras["high_order"] = bench_high_order(dev)
        extras["batched"] = bench_batched(dev)
        extras["grad"] = bench_grad(dev)
        extras["leapfrog"] = bench_leapfrog(dev)

    prof = "multipole_halo_lmax8_pot_force_N10M.md"
    traffic = profiled_value(prof, "traffic_bytes_per_launch")
    f64 = profiled_value(prof, "fp64_flops_per_launch")
    line = {
        "metric": "FP64 particles/s (Multipole lmax=8, pot+force, N=10M); batched+grad vs plain",
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "settle": {"seconds": args.settle, "steps": n_settle},
        "config": {"workload": "multipole_halo_lmax8_pot_force", "n_particles": N,
                   "lmax": 8, "radial_nodes": 50, "active_terms": opinfo["n"],
                   "table": "triaxial Dehnen gamma=1 M=1e12 a=20 axes 1:0.8:0.6",
                   "points": "Hernquist a=20 seed 0, 1% log-uniform [0.01,2000]"},
        "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": traffic,
                     "algorithmic_bytes": 56 * N,
                     "kernel_ms": k_mean, "kernel_ms_median": k_median,
                     "hbm_gbs": (traffic if traffic else 56 * N) / (k_mean * 1e-3) / 1e9,
                     "flops_per_launch": flops, "opcount": opinfo,
                     # FP64 flops the kernel executes (rocprof SQ_INSTS_VALU_{FMA,MUL,ADD,
                     # TRANS}_F64 x 64 lanes, FMA = 2) over the same launch time
                     "fp64_executed_per_launch": f64,
                     "fp64_executed_tflops": (f64 / (k_mean * 1e-3) / 1e12) if f64 else None,
                     "fp64_executed_frac": (f64 / (k_mean * 1e-3) / 1e12 / PEAK_FP64_TFLOPS)
                     if f64 else None},
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
