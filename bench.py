#!/usr/bin/env python3
"""Benchmark of the PV Newton-step linear solve on one MI355X.

Metric: Newton refactor+solve per second, batched vs plain, on 1 GPU.
A step is one refactor + L/U solve of every instance of the batch (one Newton
iteration's linear algebra).  Default workload (the batched config):
512 seeded cloud scenarios of the 40 inverters x 24 strings x 28 modules plant
(26,920 FP64 unknowns, 80,680 nonzeros under AMD ordering, zero fill), all
instances sharing one pattern.  The systems are the plant's companion-model
Jacobians stamped on the GPU at the converged Newton state of each scenario
(batched device Newton, correction form, to 1e-12 V), so values are
device-resident when the timed region starts.  Beside it on the same line:
  * plain: the single-instance GPU path on the same plant (batch 1, the
    library's default schedule), with the reference's launch structure (one
    kernel per level, one lane per chain) timed beside it;
  * dropin: the reference call sequence end to end (host assembly, copies,
    GPU refactor/solve/update) per Newton iteration;
  * roofline of the refactor+solve launch and the reference CPU GLU baseline.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--inverters", type=int, default=40)
    ap.add_argument("--strings", type=int, default=24)
    ap.add_argument("--modules", type=int, default=28)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip plain / drop-in legs")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the mixed-size (10/40/100 inverters) and 400-inverter legs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--schedule", default="hybrid:64:1",
                    help="variant:width:groups of the batched launch (level, component, "
                         "hybrid), or 'auto' for the autotuner's choice")
    ap.add_argument("--profile-newton", type=int, default=0, metavar="ITERS",
                    help="profiling mode: run ITERS replays of the batched Newton graph on "
                         "the default workload, print their timing and exit (no headline)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-cores CPU baseline (0: nproc)")
    return ap.parse_args()


SCHED_NAMES = {"level": 1, "component": 2, "hybrid": 3}


# CPU builds timed as the baseline: the reference's host flags
# (PV/Makefile:45, -O3 -march=native -funroll-loops --fast-math; built with
# -march=x86-64-v3 so the library runs on the box's host) and the same
# without --fast-math.
CPU_BUILDS = {"fast_math": True, "no_fast_math": "o3"}


def _cpu_kind(build=True):
    from oracle import oracle as O
    return "reference" if O.ref(fast=build) is not None else "port"


def _cpu_one(s, lu, x, kind, build):
    from oracle import oracle as O
    ref = kind == "reference"
    # the port (oracle restatement) has one optimized build only
    fast = build if ref else True
    O.glu_csc_sequential(s, lu, use_ref=ref, fast=fast)
    O.substitute_lu_csc(s, lu, x, use_ref=ref, fast=fast)


def _cpu_run(s, AT, bT, ks, kind, build):
    """refactor+solve of instances ks on one thread; each instance's values are
    copied into a scratch vector first (the GLU works in place)."""
    lu = np.empty(AT.shape[1])
    x = np.empty(bT.shape[1])
    for k in ks:
        np.copyto(lu, AT[k])
        np.copyto(x, bT[k])
        _cpu_one(s, lu, x, kind, build)


def host_cpu():
    """nproc (the processing units this process may use; GNU nproc honours
    the affinity mask and OMP_NUM_THREADS), nproc --all and lscpu's model."""
    def run(cmd):
        try:
            return subprocess.run(cmd, capture_output=True, text=True).stdout
        except OSError:
            return ""
    try:
        nproc = int(run(["nproc"]).strip())
    except ValueError:
        nproc = len(os.sched_getaffinity(0))
    try:
        nall = int(run(["nproc", "--all"]).strip())
    except ValueError:
        nall = os.cpu_count() or nproc
    model = [l.split(":", 1)[1].strip() for l in run(["lscpu"]).splitlines()
             if l.startswith("Model name")]
    return {"nproc": nproc, "nproc_all": nall, "model": model[0] if model else "unknown"}


def cpu_baseline(plan_arrays, A_host, b_host, threads=1, budget_s=12.0, build=True):
    """Reference CPU GLU (GLU_CSC_CPU_sequential + Substitute_LU_Vector_b_CSC_CPU
    compiled from the reference's own sources; build: True = its host flags
    with --fast-math, "o3" = without) on a bounded sample of the same
    instances, on `threads` threads (the foreign calls release the
    interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    kind = _cpu_kind(build)
    s = O.CsrCsc(*plan_arrays)
    B = A_host.shape[1]
    AT = np.ascontiguousarray(A_host.T)
    bT = np.ascontiguousarray(b_host.T)
    _cpu_run(s, AT, bT, [k % B for k in range(5)], kind, build)   # warm-up
    t0 = time.perf_counter()
    _cpu_run(s, AT, bT, [k % B for k in range(20)], kind, build)
    one = max((time.perf_counter() - t0) / 20, 1e-6)
    count = int(min(max(threads, budget_s * threads / one), 400 * B))
    ks = [k % B for k in range(count)]
    t0 = time.perf_counter()
    if threads == 1:
        _cpu_run(s, AT, bT, ks, kind, build)
    else:
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda j: _cpu_run(s, AT, bT, ks[j::threads], kind, build),
                        range(threads)))
    dt = time.perf_counter() - t0
    flags = ("-O3 -march=x86-64-v3 -funroll-loops" +
             (" --fast-math" if build is True else ""))
    return {"value": count / dt, "unit": "refactor+solve/s", "cores": threads, "kind": kind,
            "build": flags if kind == "reference" else "oracle port, -O3 -march=x86-64-v3",
            "sample": f"{count} instances of the batch (refactor+solve each, values copied "
                      f"in per instance), {dt:.1f} s, {threads} thread"
                      + ("s" if threads > 1 else "")}


def _time_ms(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _oracle_newton(pv, plant):
    """The oracle Newton of a plant (correction form, FP64 CPU)."""
    from oracle import oracle as O
    pos, neg, rsh = plant.modules()
    a, b, g = plant.cables()
    rp, ci = plant.pattern()
    P, Prow = pv.reorder(plant.n, rp, ci)
    return O.PVNewton(plant.n, pos, neg, rsh, a, b, g, plant.buses(), P, Prow,
                      form="correction")


def newton_error(on, plant, x, v, G, ks, cold=False):
    """max |x - x_oracle| (V) over instances ks: the oracle Newton iterated on
    its own to 1e-12 V, from the library's cold start (cold=True) or from the
    device's converged element voltages (it then runs until its own last
    update is below 1e-12 V, i.e. to its own fixed point)."""
    worst = 0.0
    for k in ks:
        v0 = plant.v0() if cold else v[:, k]
        xo, _, _ = on.solve(9e-3 * G[:, k], v0, ladder=False, tol=1e-12)
        worst = max(worst, float(np.abs(x[:, k] - xo).max()))
    return worst


def _spread(its):
    its = np.asarray(its)
    return [int(its.min()), int(np.median(its)), int(its.max())]


def config_legs(torch, pv, dev):
    """Secondary workloads, reported beside the headline value:
    cfg0: the 4 x 10 x 20 plant, drop-in Newton (the reference's call
    sequence) against the oracle;
    cfg1 window: the 40 x 24 x 28 plant, one scenario, 200 one-second time
    steps, each a warm-started Newton solve on the single-instance device path;
    mixed (cfg3): 128 instances round-robin over 10 / 40 / 100-inverter
    plants, batched Newton per plant, one multi-pattern launch for
    refactor+solve and one for the adjoint solve;
    large (cfg4): the 400-inverter plant (269,200 unknowns), 32 scenarios,
    50 one-second steps of batched Newton with the energy adjoint of each."""
    out, acc = {}, {}
    # ------------------------------------------------------------ cfg0
    p0 = pv.PV(4, 10, 20, seed=0)
    g0 = p0.irradiance(-1)
    x0, v0, st0 = p0.newton_dropin(g0, tol=1e-12, ladder=True)
    on0 = _oracle_newton(pv, p0)
    acc["cfg0_4x10x20"] = {
        "max_abs_dx_vs_oracle_V": newton_error(on0, p0, x0[:, None], v0[:, None],
                                               g0[:, None], [0], cold=True),
        "instances": 1, "checked": 1, "oracle_start": "cold (ladder off, 1e-12 V)",
        "path": "drop-in (reference call sequence), correction form, ladder",
        "iterations": int(st0[0])}
    del p0, on0
    # ------------------------------------------------------------ cfg1
    p1 = pv.PV(40, 24, 28, seed=0)
    steps = 200
    v = None
    its, xs = [], {}
    t0 = time.perf_counter()
    loop_ms = 0.0
    for t in range(steps):
        G = p1.irradiance(0, float(t))[:, None].copy()
        x, v, p, st = p1.newton_batched(G, v=v, tol=1e-12, ladder=False)
        its.append(int(st[0]))
        loop_ms += float(st[1])
        if t in (0, steps - 1):
            xs[t] = (x, v, G)
    wall = time.perf_counter() - t0
    on1 = _oracle_newton(pv, p1)
    e_first = newton_error(on1, p1, *xs[0], [0], cold=True)
    e_last = newton_error(on1, p1, *xs[steps - 1], [0])
    out["window_cfg1_40x24x28"] = {
        "steps": steps, "wall_s": round(wall, 2), "ms_per_step": round(wall / steps * 1e3, 2),
        "newton_loop_ms_per_step": round(loop_ms / steps, 2),
        "iterations_per_step": _spread(its), "iterations_total": int(sum(its)),
        "note": "one cloud scenario, 200 one-second steps, each warm-started from the "
                "previous step's state; single-instance device path (batch 1: device "
                "stamping, COMPONENT refactor+solve, graph replays); host cloud field per step"}
    acc["cfg1_40x24x28_window"] = {
        "max_abs_dx_vs_oracle_V": max(e_first, e_last), "instances": steps, "checked": 2,
        "oracle_start": "cold at t = 0; from the device state at t = 199",
        "path": "single-instance device Newton, correction form"}
    del p1, on1
    # ------------------------------------------------------------ cfg3
    plants = [pv.PV(I, 24, 28, seed=3) for I in (10, 40, 100)]
    batches = [len(range(g, 128, 3)) for g in range(3)]
    As, bs, gs, lus, xs_, ys, nbytes, abytes = [], [], [], [], [], [], 0, 0
    worst, floor, nits = 0.0, 0, []
    for g, (p, B) in enumerate(zip(plants, batches)):
        plan = p.plan()
        G = np.stack([p.irradiance(s, 0.0) for s in range(g, 128, 3)], axis=1)
        x, v, _, _ = p.newton_batched(G, tol=1e-12, ladder=False)
        it, last = p.newton_info(B)
        floor += int((last > 1e-12).sum())
        nits.extend(it.tolist())
        worst = max(worst, newton_error(_oracle_newton(pv, p), p, x, v, G, [B // 2]))
        A = torch.empty((plan.nnz, B), dtype=torch.float64, device=dev)
        b = torch.empty((p.n, B), dtype=torch.float64, device=dev)
        p.stamp(B, torch.from_numpy(G).to(dev), torch.from_numpy(v).to(dev), A, b, -1.0)
        As.append(A), bs.append(b), gs.append(b.clone())
        lus.append(torch.empty_like(A)), xs_.append(torch.empty_like(b))
        ys.append(torch.empty_like(b))
        nbytes += B * (24 * plan.nnz + 16 * p.n) + 4 * plan.nnz + 8 * p.n
        abytes += B * (8 * plan.nnz + 16 * p.n)
    m = pv.Multi([p.plan() for p in plants], batches)
    t_rs = _time_ms(torch, lambda: m.refactor_solve(As, lus, bs, xs_), 20)
    t_adj = _time_ms(torch, lambda: m.solve_transpose(lus, gs, ys), 20)
    out["mixed_10_40_100"] = {
        "instances": 128, "refactor_solve_ms": round(t_rs, 4),
        "refactor_solve_per_s": round(128 / t_rs * 1e3, 1),
        "refactor_solve_GBps": round(nbytes / t_rs / 1e6, 1),
        "adjoint_ms": round(t_adj, 4), "adjoint_GBps": round(abytes / t_adj / 1e6, 1),
        "note": "systems stamped at each instance's converged Newton state"}
    acc["cfg3_mixed_10_40_100"] = {
        "max_abs_dx_vs_oracle_V": worst, "instances": 128, "checked": 3,
        "noise_floor_stops": floor, "iterations_per_instance": _spread(nits),
        "oracle_start": "from the device state (one instance per plant size)",
        "path": "batched device Newton per plant, correction form"}
    del As, bs, gs, lus, xs_, ys, m, plants
    # ------------------------------------------------------------ cfg4
    big = pv.PV(400, 24, 28, seed=4)
    B, steps = 32, 50
    gdev = torch.empty((big.nmod, B), dtype=torch.float64, device=dev)
    v = None
    energy = np.zeros(B)
    gnorm2 = 0.0
    its, floor, step_ms = [], 0, []
    first = None
    t0 = time.perf_counter()
    for t in range(steps):
        ts = time.perf_counter()
        big.irradiance_device(list(range(B)), [float(t)] * B, gdev)
        G = gdev.cpu().numpy()
        # forward Newton (warm from the previous second) + energy adjoint:
        # E = sum_t P_t * 1 s, so dE/dG_{m,t} = dP_t/dG_{m,t}
        p, dG, dgc, v = big.gradients(G, v=v, tol=1e-12)
        it, last = big.newton_info(B)
        its.append(_spread(it))
        floor += int((last > 1e-12).sum())
        energy += p * 1.0
        gnorm2 += float(np.sum(dG * dG))
        step_ms.append((time.perf_counter() - ts) * 1e3)
        if t == 0:
            first = (G.copy(), v.copy())
    wall = time.perf_counter() - t0
    plan = big.plan()
    Gl, vl = first
    A = torch.empty((plan.nnz, B), dtype=torch.float64, device=dev)
    b = torch.empty((big.n, B), dtype=torch.float64, device=dev)
    big.stamp(B, torch.from_numpy(Gl).to(dev), torch.from_numpy(vl).to(dev), A, b, -1.0)
    lu, xx = torch.empty_like(A), torch.empty_like(b)
    t_rs = _time_ms(torch, lambda: plan.refactor_solve(A, lu, b, xx, B), 20)
    byts = B * (24 * plan.nnz + 16 * big.n) + 4 * plan.nnz + 8 * big.n
    # accuracy at t = 0 (two scenarios)
    x0, v0, _, _ = big.newton_batched(Gl, v=vl, tol=1e-12, ladder=False)
    e4 = newton_error(_oracle_newton(pv, big), big, x0, v0, Gl, [3, 30])
    out["large_400x24x28"] = {
        "n": big.n, "scenarios": B, "steps": steps, "window_wall_s": round(wall, 2),
        "ms_per_step": round(float(np.median(step_ms)), 1),
        "iterations_per_step_min_median_max": its[:3] + ["..."] + its[-2:],
        "energy_MWh_per_scenario_mean": round(float(energy.mean()) / 3.6e9, 6),
        "dE_dG_norm_Ws_per_Wm2": round(gnorm2 ** 0.5, 3),
        "refactor_solve_ms": round(t_rs, 4), "refactor_solve_GBps": round(byts / t_rs / 1e6, 1),
        "note": "per step: device cloud field (32 x 268,800 modules), batched Newton warm-"
                "started from the previous second to 1e-12 V, energy adjoint dP_t/dG_{m,t} "
                "for every module (transposed solves with the step's factors); "
                "ms_per_step includes the host copies of G, v and dE/dG"}
    acc["cfg4_400x24x28"] = {
        "max_abs_dx_vs_oracle_V": e4, "instances": B * steps, "checked": 2,
        "noise_floor_stops": floor, "oracle_start": "from the device state (t = 0)",
        "path": "batched device Newton, correction form"}
    return out, acc


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the path shards by instances with no "
              "exchange inside Newton, and this build runs on a single MI355X. "
              "Run with --gpus 1.", file=sys.stderr)
        return 2
    import torch
    import phdthesis_albertojimenezruiz_amd as pv

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    B = args.batch
    plant = pv.PV(args.inverters, args.strings, args.modules, seed=args.seed)
    plan = plant.plan()
    n, nnz = plant.n, plan.nnz
    print(f"[bench] plant {args.inverters}x{args.strings}x{args.modules} n={n} nnz={nnz} "
          f"batch={B} ncomp={plan.info('ncomp')} levels lu/l/u="
          f"{plan.info('nlev_lu')}/{plan.info('nlev_l')}/{plan.info('nlev_u')}", flush=True)

    # Converged operating point of each scenario (batched device Newton)
    G = np.stack([plant.irradiance(s, 0.0) for s in range(B)], axis=1)
    if args.profile_newton > 0:
        # a few replays from the cold start: every instance active, the
        # kernels of one Newton iteration as the solve runs them, on the
        # schedule given (counter collection serializes the autotuner's
        # timing runs, so its pick under a counter pass is not the loop's)
        if args.schedule != "auto":
            v, w, g = args.schedule.split(":")
            plan.set_schedule(SCHED_NAMES[v], int(w))
            plan.set_stream_groups(int(g))
        plant.newton_batched(G, tol=1e-12, max_iters=2, ladder=False, raise_noconv=False)
        t0 = time.perf_counter()
        plant.newton_batched(G, tol=1e-12, max_iters=args.profile_newton, ladder=False,
                             raise_noconv=False)
        dt = time.perf_counter() - t0
        print(f"[bench] profile-newton: {args.profile_newton} replays of the batched Newton "
              f"graph ({B} x {args.inverters}x{args.strings}x{args.modules}), "
              f"{dt / args.profile_newton * 1e3:.3f} ms per replay incl. setup", flush=True)
        return 0
    t0 = time.perf_counter()
    state = "converged Newton state (max|dV| <= 1e-12 V, or <= 1e-11 V at the noise floor)"
    newton = None
    try:
        x_n, v_n, power, nstats = plant.newton_batched(G, tol=1e-12, ladder=False)
        t_newton = time.perf_counter() - t0
        its, last = plant.newton_info(B)
        newton = {"replays": int(nstats[0]), "ms_per_iteration": round(float(nstats[2]), 4),
                  "iterations_per_instance": _spread(its),
                  "noise_floor_stops": int((last > 1e-12).sum()),
                  "max_final_update_V": float(last.max())}
        print(f"[bench] batched Newton: {nstats[0]:.0f} iterations, {nstats[2]:.3f} ms/iter, "
              f"max|dV| {nstats[3]:.2e} V, total power {power.sum() / 1e6:.2f} MW "
              f"({t_newton:.1f} s incl. setup)", flush=True)
    except pv.PvgluError as e:
        print(f"[bench] batched Newton failed ({e}); timing the systems at the initial "
              f"state instead", flush=True)
        v_n = plant.v0(B)
        state = "initial Newton state (V = 600/M per module)"
    if args.schedule == "auto":
        plan.autotune(B, 5)
    else:
        v, w, g = args.schedule.split(":")
        plan.set_schedule(SCHED_NAMES[v], int(w))
        plan.set_stream_groups(int(g))
        # the cache policy by its rule (the batched Newton's autotune may have
        # left its own pick on the plan): streaming for this 1.2 GB batch
        plan.set_cache_policy(pv.CACHE_AUTO)
    sched = (*plan.get_schedule(), plan.get_stream_groups())
    sched_name = {1: "level", 2: "component", 3: "hybrid"}[sched[0]] + f":{sched[1]}:{sched[2]}"

    dG = torch.from_numpy(G).to(dev)
    dV = torch.from_numpy(v_n).to(dev)
    A = torch.empty((nnz, B), dtype=torch.float64, device=dev)
    b = torch.empty((n, B), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    plan.status_reset(stream)
    plant.stamp(B, dG, dV, A, b, -1.0, stream)
    LU = torch.empty_like(A)
    X = torch.empty_like(b)
    torch.cuda.synchronize()

    # ----------------------------------------------------------- batched
    for _ in range(args.warmup):
        plan.refactor_solve(A, LU, b, X, B, stream)
    torch.cuda.synchronize()
    # the timed region: K launches back to back between two events on the
    # launch stream (no instrumentation between the launches)
    t_all0 = torch.cuda.Event(enable_timing=True)
    t_all1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_all0.record(stream)
    for _ in range(args.steps):
        plan.refactor_solve(A, LU, b, X, B, stream)
    t_all1.record(stream)
    torch.cuda.synchronize()
    total_ms = t_all0.elapsed_time(t_all1)
    # the roofline's launch duration: a second pass of K launches, each
    # between its own pair of events on the launch stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    for e0, e1 in ev:
        e0.record(stream)
        plan.refactor_solve(A, LU, b, X, B, stream)
        e1.record(stream)
    torch.cuda.synchronize()
    launch_ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    avg_launch_ms = float(np.mean(launch_ms))
    ms_per_step = total_ms / args.steps
    value = B * args.steps / (total_ms / 1e3)
    code = plan.status()[0]
    if code:
        raise RuntimeError(f"device status {code} after the timed region")

    # accuracy spot check against the oracle (bit-exact expected)
    from oracle import oracle as O
    arrays = [plan.array(k) for k in ("csr_rowptr", "csr_colidx", "csc_colptr", "csc_rowidx",
                                      "csr2csc", "csc2csr", "csr_diag", "csc_diag")]
    plan_arrays = (n, *arrays)
    s = O.CsrCsc(*plan_arrays)
    A_host = A.cpu().numpy()
    b_host = b.cpu().numpy()
    X_host = X.cpu().numpy()
    max_dx = 0.0
    for k in (0, B // 2, B - 1):
        lu = np.ascontiguousarray(A_host[:, k])
        xx = np.ascontiguousarray(b_host[:, k])
        O.glu_csc_sequential(s, lu)
        O.substitute_lu_csc(s, lu, xx)
        max_dx = max(max_dx, float(np.abs(xx - X_host[:, k]).max()))

    bytes_launch = B * (24 * nnz + 16 * n) + (4 * nnz + 8 * n)
    achieved = bytes_launch / (avg_launch_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "kernel": f"refactor+solve launch (schedule {sched_name})",
                "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg_launch_ms, 5)}
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        try:
            t = json.load(open(tf))
            if t.get("batch") == B and t.get("n") == n and t.get("schedule") == sched_name:
                roofline["traffic"] = t.get("bytes_per_launch")
                roofline["traffic_source"] = (
                    "calibrated constant from profiles/traffic.json: two rocprofv3 --pmc "
                    "passes (FETCH_SIZE, WRITE_SIZE) of this workload and schedule, scaled "
                    "by the counters of a 1 GiB device copy; not measured in this run")
        except (OSError, ValueError):
            pass
    print(f"[bench] batched: {ms_per_step:.4f} ms/step ({avg_launch_ms:.4f} ms/launch), "
          f"{value:.4g} refactor+solve/s, {achieved:.0f} GB/s algorithmic, "
          f"max|dx| vs oracle {max_dx:.1e}", flush=True)

    out = {"metric": "Newton refactor+solve/sec (batched vs plain, 1 GPU) at <=1e-10 V; "
                     "% HBM peak",
           "value": value, "unit": "refactor+solve/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded plant + cloud scenarios, companion Jacobians at the "
                   + state + ")",
           "config": {"workload": f"batched: {B} cloud scenarios x {args.inverters}x"
                                  f"{args.strings}x{args.modules} plant, shared pattern",
                      "batch": B, "n": n, "nnz_lu": nnz, "ordering": "AMD",
                      "schedule": sched_name,
                      "newton_tol_V": 1e-10, "max_abs_dx_vs_oracle": max_dx},
           "roofline": roofline}
    if newton is not None:
        # per-config Newton accuracy: node voltages against the FP64 oracle Newton
        on = _oracle_newton(pv, plant)
        ks = [0, B // 3, 2 * B // 3, B - 1]
        out["newton_accuracy"] = {"cfg2_40x24x28_b512": dict(
            newton, max_abs_dx_vs_oracle_V=newton_error(on, plant, x_n, v_n, G, ks),
            instances=B, checked=len(ks), oracle_start="from the device state",
            path="batched device Newton (hipGraph per iteration), correction form")}
        del on
    else:
        out["newton_accuracy"] = {}

    # ------------------------------------------------------------- plain
    if not args.no_plain:
        p1 = pv.PV(args.inverters, args.strings, args.modules, seed=args.seed)
        plan1 = p1.plan()
        A1 = A[:, :1].contiguous()
        b1 = b[:, :1].contiguous()
        LU1 = torch.empty_like(A1)
        X1 = torch.empty_like(b1)

        def time_single():
            for _ in range(args.warmup):
                plan1.refactor_solve(A1, LU1, b1, X1, 1, stream)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            reps = max(args.steps, 20)
            e0.record(stream)
            for _ in range(reps):
                plan1.refactor_solve(A1, LU1, b1, X1, 1, stream)
            e1.record(stream)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        # the library's single-instance path (AUTO: one COMPONENT launch) and
        # the reference's launch structure (one kernel per level, one lane
        # per chain) beside it
        plain_ms = time_single()
        sv, sw = plan1.get_schedule()
        plan1.set_schedule(pv.SCHED_LEVEL, 1)
        level_ms = time_single()
        plan1.set_schedule(sv, sw)
        plain = 1e3 / plain_ms
        out["plain"] = {"value": plain, "unit": "refactor+solve/s", "ms_per_solve": plain_ms,
                        "schedule": "library default at batch 1 (COMPONENT, one launch)",
                        "reference_structure_ms_per_solve": level_ms,
                        "reference_structure": "one kernel per level, 1 lane per chain"}
        out["batched_over_plain"] = value / plain
        # drop-in: the reference call sequence per Newton iteration
        xd, vd, st = p1.newton_dropin(G[:, 0], v=v_n[:, 0], fixed_iters=5)
        out["dropin"] = {"ms_per_newton_iteration": st[3] / st[0],
                         "host_assembly_ms": st[1] / st[0], "device_ms": st[2] / st[0],
                         "note": "host assembly + H2D/D2H + GPU refactor/solve/update"}
        print(f"[bench] plain: {plain_ms:.4f} ms/solve ({plain:.4g}/s; reference launch "
              f"structure {level_ms:.4f} ms); batched/plain {value / plain:.1f}x; drop-in "
              f"{st[3] / st[0]:.2f} ms/iteration", flush=True)

    if not args.no_cpu:
        host = host_cpu()
        threads = args.cpu_threads if args.cpu_threads > 0 else host["nproc"]
        legs = {}
        for name, build in CPU_BUILDS.items():
            cb = cpu_baseline(plan_arrays, A_host, b_host, build=build, budget_s=10.0)
            ct = cpu_baseline(plan_arrays, A_host, b_host, threads=threads, budget_s=6.0,
                              build=build)
            legs[name] = (cb, ct)
            print(f"[bench] cpu baseline {name} ({cb['kind']}): {cb['value']:.4g}/s on 1 "
                  f"thread, {ct['value']:.4g}/s on {threads} threads ({host['model']}, "
                  f"nproc {host['nproc']}, nproc --all {host['nproc_all']})", flush=True)
        cb, ct = legs["fast_math"]
        out["cpu_baseline"] = dict(cb, cpu=host["model"], nproc=host["nproc"],
                                   nproc_all=host["nproc_all"])
        out["batched_over_cpu"] = value / cb["value"]
        out["cpu_baseline_all_cores"] = dict(ct, cpu=host["model"])
        out["batched_over_cpu_all_cores"] = value / ct["value"]
        nb, nt = legs["no_fast_math"]
        out["cpu_baseline_no_fast_math"] = dict(nb, cpu=host["model"])
        out["cpu_baseline_no_fast_math_all_cores"] = dict(nt, cpu=host["model"])
        out["batched_over_cpu_no_fast_math"] = value / nb["value"]

    if not args.no_configs:
        out["configs"], acc = config_legs(torch, pv, dev)
        out["newton_accuracy"].update(acc)
        for k, v in out["configs"].items():
            print(f"[bench] {k}: {v}", flush=True)
        for k, v in acc.items():
            print(f"[bench] accuracy {k}: {v}", flush=True)

    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = 2 * src.numel() * 8

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
