"""Benchmark of the TDVP time step on one MI355X.

Default workload: the batched configuration -- 64 uneven spin-boson instances (Ohmic baths,
alpha in [0.05, 1], wc in {1, 2, 5}, beta in {1, 10, inf}, 32-128 modes, boson dims 8-20,
bond dims 16-48, dt 0.05, complex128) advanced together by the persistent sweep kernel.

One "step" is one TDVP time step of every instance in the batch.  The timed region covers
exactly --steps batched steps after --warmup untimed ones (the first step of a run carries
the reference's prepare pass and six-sub-step ladder, so it is always part of the warmup).
`value` is instance-steps per second of the whole batch.  Besides it the JSON line carries:
  roofline      algorithmic FP64 flops of the timed steps (H applications counted on the
                device, per-node cost from the compiled term tables) / kernel time, against
                the MI355X FP64 matrix peak (78.6 TFLOP/s); traffic is null (HBM bytes come
                from rocprofv3 --pmc runs, which this script does not make)
  plain_gpu     the plain per-instance path: a sample of the instances (the slowest, the
                fastest and the median ones by their device time in the batch) restored from
                checkpoints taken before the timed steps and stepped alone, each for the same
                step as in the batch; the other instances' alone times are estimated from their
                batch device time times the sample's median alone/batch ratio
  cpu_baseline  the oracle (C++ CPU restatement, 1 thread) on a bounded sample of instances
                (lightest first, then one heavier one if the budget allows), a regular step
                each after an untimed first step, scaled to the whole batch by flop count;
                the same steps on the GPU give check.oracle_max_dsigma_z
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TDVP steps/s + FP64-MFMA roofline %, batched vs plain GPU vs CPU, <sz> err<=1e-10"
PEAK_FP64_TFLOPS = 78.6


def make(lib, inst):
    from ttns_lib_amd.capi import SIGMA, Integrator, sb_problem

    tree, sop = sb_problem(lib, inst.model)
    ig = Integrator(lib, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-12)
    ig.set_sb_initial_state(inst.model)
    ig.set_observable(0, SIGMA["z"])
    return ig


def workload(name, lib=None):
    from ttns_lib_amd import workloads as W

    if name == "c3":
        return W.c3_instances(), ("64 uneven spin-boson instances, Ohmic alpha U[0.05,1], wc {1,2,5}, "
                                  "beta {1,10,inf}, 32-128 modes, boson 8-20, bond 16-48, dt 0.05")
    if name == "c0":
        return [W.c0_instance()], "16 modes, boson 10, bond 8, Ohmic alpha 0.1 wc 5 T=0, dt 0.05"
    if name == "c0x64":
        insts = []
        for s in range(64):
            i = W.c0_instance()
            i.model.seed = s
            insts.append(i)
        return insts, "64 copies of the 16-mode C0 instance (different random unoccupied functions)"
    if name == "c1":
        return [W.c1_instance()], "Debye alpha 0.5 wc 2 beta 10, 128 modes, boson 16, bond 32 (fixed), dt 0.025"
    if name == "c4":
        return [W.c4_instance()], "Ohmic alpha 0.2 wc 5 T=0, 64 modes, boson 12, bond 24, dt 0.05"
    if name in ("refruns", "refruns256"):
        n = 64 if name == "refruns" else 256
        return W.reference_runs(lib, n), (f"{n} instances of the reference's five 512-mode example runs (exact "
                                          "baths, driver mode blocks, bonds 4-8, dt 0.005), seeds varied")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(lib, insts, flops_per_step, budget_s):
    """The oracle (1 thread) on a bounded sample: the lightest instances first, then the lightest
    instance above the median flop count if the measured rate says it fits the remaining
    budget; an untimed first step, then one timed regular step each.  The CPU's flop rate over
    the sample scales to the whole batch.  The same two steps on the GPU check the oracle's
    <sigma_z> against the HIP path's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from ttns_lib_amd.capi import SIGMA, Integrator, Library, sb_problem

    so = os.path.join(ROOT, "oracle", "_build", "libttn_oracle.so")
    if not os.path.exists(so):
        return None, None
    orc = Library(so, "orc_")
    order = [int(i) for i in np.argsort(flops_per_step)]
    heavy = next(i for i in order if flops_per_step[i] >= np.median(flops_per_step))
    used, per, dsz = [], [], 0.0
    t_cpu = f_cpu = 0.0
    t_start = time.perf_counter()

    def one(i):
        nonlocal t_cpu, f_cpu, dsz
        inst = insts[i]
        tree, sop = sb_problem(orc, inst.model)
        ig = Integrator(orc, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-12)
        ig.set_sb_initial_state(inst.model)
        ig.step(1)  # first step (ladder) untimed
        t0 = time.perf_counter()
        ig.step(1)
        dt = time.perf_counter() - t0
        g = make(lib, inst)
        g.step(2)
        dsz = max(dsz, abs(g.expect(0, SIGMA["z"]).real - ig.expect(0, SIGMA["z"]).real))
        g.destroy()
        used.append(i)
        per.append(flops_per_step[i] / dt / 1e9)
        t_cpu += dt
        f_cpu += flops_per_step[i]

    for i in order:
        if i == heavy or (used and time.perf_counter() - t_start > budget_s / 2):
            break
        one(i)
    # a single instance (or a batch whose lightest instance is already the median) is timed
    # as it is; otherwise the heavy one joins when the rate so far says it fits the budget
    if not used or 2.3 * flops_per_step[heavy] / (f_cpu / t_cpu) < budget_s - (time.perf_counter() - t_start):
        one(heavy)
    rate = f_cpu / t_cpu
    total_time = float(np.sum(flops_per_step)) / rate  # one batched step on one core
    return {
        "value": len(insts) / total_time,
        "unit": "instance-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle, 1 thread, one regular step (after an untimed first step) of instances {used} "
                  f"({', '.join(f'{flops_per_step[i] / 1e9:.1f}' for i in used)} GFLOP per step; "
                  f"{', '.join(f'{r:.2f}' for r in per)} GFLOP/s); {rate / 1e9:.3g} GFLOP/s over the sample "
                  f"scales the batch's {np.sum(flops_per_step) / 1e12:.2f} TFLOP per step (extrapolated)",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
    }, dsz


def _cpu_worker(so, k, inst, barrier, queue):
    """One oracle process of the all-cores baseline: first step untimed, then (after every
    process has reached the barrier) one timed regular step."""
    sys.path.insert(0, ROOT)
    from ttns_lib_amd.capi import Integrator, Library, sb_problem

    orc = Library(so, "orc_")
    tree, sop = sb_problem(orc, inst.model)
    ig = Integrator(orc, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-12)
    ig.set_sb_initial_state(inst.model)
    ig.step(1)
    barrier.wait(timeout=600)
    t0 = time.perf_counter()
    ig.step(1)
    queue.put((k, time.perf_counter() - t0))


def cpu_baseline_all_cores(insts, flops_per_step, cores):
    """The oracle on `cores` processes at once (one instance each, the cheapest ones, as the
    reference's ensemble runner does with a process pool): the sum of the processes' flop rates
    during their concurrent regular step, scaled to the batch by per-instance flop count."""
    import multiprocessing as mp

    so = os.path.join(ROOT, "oracle", "_build", "libttn_oracle.so")
    if not os.path.exists(so):
        return None
    ctx = mp.get_context("spawn")
    order = [int(i) for i in np.argsort(flops_per_step)[:cores]]
    barrier, queue = ctx.Barrier(len(order)), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=(so, i, insts[i], barrier, queue)) for i in order]
    for p in procs:
        p.start()
    times = dict(queue.get(timeout=900) for _ in procs)
    for p in procs:
        p.join(timeout=60)
    rate = float(sum(flops_per_step[i] / times[i] for i in order))  # flop/s of all processes together
    total_time = float(np.sum(flops_per_step)) / rate
    return {
        "value": len(insts) / total_time,
        "unit": "instance-steps/s",
        "cores": len(order),
        "kind": "port",
        "sample": f"oracle, {len(order)} processes at once, one regular step each of instances {order}; "
                  f"{rate / 1e9:.3g} GFLOP/s together, scaled to the batch by per-instance flop count",
    }


def report(insts, igs, teams, dflops, p0, steps, s0=None):
    """Per-instance device phase times (ttn_profile, seconds per step), slowest first; for the
    slowest instance also the H applications per node shape (node and bond propagations)."""
    rows = []
    for k, ig in enumerate(igs):
        p1 = ig.profile()
        d = {n: (p1[n] - p0[k][n]) / (1 if n == "n_small" else 1e6) / steps for n in p1}
        rows.append((d["total"], k, d))
    rows.sort(reverse=True)
    for tot, k, d in rows:
        m = insts[k].model
        gf = dflops[k] / steps / 1e9
        d0 = f"{m.boson_dim[0]:2d}" if m.boson_dim is not None else "bl"  # "bl": leaves hold mode blocks
        print(f"[report] inst {k:2d} modes {len(m.wk):3d} d {d0} r {m.nspf:2d} team {teams[k]:2d} "
              f"{gf:8.1f} GFLOP {tot:6.2f} s {gf / max(tot, 1e-9):6.1f} GFLOP/s | "
              + " ".join(f"{n} {v:.2f}" for n, v in d.items() if n != "total"), file=sys.stderr, flush=True)
    if s0 is not None:
        k = rows[0][1]
        ig = igs[k]
        st = (ig.node_stats() - s0[k]) / steps
        groups = {}
        for n, shp in enumerate(ig.shapes):
            g = groups.setdefault(shp, np.zeros(9))
            g += np.array([1, *st[n]])
        for shp, g in sorted(groups.items(), key=lambda kv: -kv[1][1]):
            print(f"[report] inst {k} nodes {shp[0]}x{shp[1]} x{int(g[0])}: node H-apps {g[1]:.0f} "
                  f"(mean Krylov size {g[2] / max(g[1], 1):.1f}, {g[5] / 1e8:.2f} s, "
                  f"{g[5] / 100 / max(g[1], 1):.1f} us per iteration of which H-action {g[7] / 100 / max(g[1], 1):.1f}), "
                  f"bond H-apps {g[3]:.0f} (mean size {g[4] / max(g[3], 1):.1f}, {g[6] / 1e8:.2f} s, "
                  f"{g[6] / 100 / max(g[3], 1):.1f} us per iteration of which H-action {g[8] / 100 / max(g[3], 1):.1f}) "
                  "per step", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_c2(args):
    """Contraction microbench: one 64 x 64 x 64 node, 40 product terms of random Hermitian 64 x 64
    operators (N(0,1) re/im, (X + X^H)/2, seed 0) on child 0, child 1 and the mean field; the
    H-action, the mean-field update of child 0 and a Krylov propagation exp(-i tau H) with
    Krylov dimension 16, on the whole GPU (ttn_node_*).  One "step" is one H-action."""
    import ttns_lib_amd
    from ttns_lib_amd.capi import Node

    lib = ttns_lib_amd.load_library()
    rng = np.random.default_rng(0)
    d, nt = 64, 40

    def herm(n, scale=1.0):
        X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        return scale * 0.5 * (X + X.conj().T)

    op0 = np.array([herm(d) for _ in range(nt)])
    op1 = np.array([herm(d) for _ in range(nt)])
    mf = np.array([herm(d) for _ in range(nt)])
    A = rng.standard_normal((d * d, d)) + 1j * rng.standard_normal((d * d, d))
    # the fixed 64 x 64 tile first, then the autotuner's choice per GEMM shape
    lib.set_autotune(False)
    node0 = Node(lib, op0, op1, mf)
    node0.set(A)
    node0.heff_timed(max(1, args.warmup))
    ms0, _ = node0.heff_timed(args.steps)
    node0.destroy()
    lib.set_autotune(True)
    # the two forms of the H-action's mode-0 and mean-field stages, then the autotuner's pick
    forms, mf_forms = {}, {}
    for form, name in ((1, "staged"), (2, "fused")):
        lib.set_heff_form(form)
        nf = Node(lib, op0, op1, mf)
        nf.set(A)
        nf.heff_timed(max(1, args.warmup))
        forms[name] = nf.heff_timed(args.steps)[0] / args.steps
        nf.mf(0, reps=1)
        nf.mf(0, reps=args.steps)
        mf_forms[name] = nf.last()[0] / args.steps
        nf.destroy()
    lib.set_heff_form(0)
    fused = forms["fused"] < forms["staged"]
    node = Node(lib, op0, op1, mf)
    node.set(A)
    node.heff_timed(max(1, args.warmup))
    ms, fl = node.heff_timed(args.steps)
    heff_ms = ms / args.steps
    from ttns_lib_amd.capi import GEMM_TILES

    tuned = [{"M": e[0], "N": e[1], "K": e[2], "batch": e[3], "tile": list(GEMM_TILES[e[5]]), "k_split": e[6],
              "ms": round(e[7], 4)} for e in lib.autotune_table()]
    node.mf(0, reps=1)
    node.mf(0, reps=args.steps)
    mf_ms, mf_fl = node.last()
    # Krylov propagation with dimension 16: operators scaled by 1/d, tau raised until the
    # propagation uses the whole 16-vector space
    nodek = Node(lib, op0 / d, op1 / d, mf / d)
    A1 = A / np.linalg.norm(A)
    for tau in (0.05, 0.1, 0.2, 0.4, 0.8):
        nodek.set(A1)
        na = nodek.expv(tau, krylov_dim=16, tol=1e-12)
        kms, kfl = nodek.last()
        if na >= 16:
            break
    achieved = fl / (ms / 1e3) / 1e12
    out = {
        "metric": METRIC,
        "value": 1e3 / heff_ms,
        "unit": "H-actions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": heff_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random Hermitian operators and node tensor)",
        "config": {"workload": "c2: node 64x64x64, 40 product terms of random Hermitian 64x64 operators",
                   "terms": nt},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": ("ttnd::zgemm_kernel (mode-1 stage) + ttnd::zheff_fused_kernel" if fused
                                else "ttnd::zgemm_kernel (three GEMM stages of the H-action)"),
                     "flops_per_heff": fl / args.steps},
        "autotune": {"heff_ms_fixed_64x64": ms0 / args.steps, "heff_ms_tuned": heff_ms, "table": tuned,
                     "heff_ms_by_form": forms},
        "mean_field": {"ms": mf_ms / args.steps, "tflops": mf_fl / (mf_ms / 1e3) / 1e12, "ms_by_form": mf_forms},
        "krylov16": {"ms": kms, "tau": tau, "h_applications": na, "tflops": kfl / (kms / 1e3) / 1e12,
                     "ms_per_iteration": kms / max(na, 1)},
    }
    print(json.dumps(out))


def bench_c4grad(args):
    """Gradient workload: d<sigma_z(T)>/d(g_k, w_k) for the 64-mode Ohmic bath (alpha 0.2,
    wc 5, T = 0, boson dim 12, bond dim 24, dt 0.05) by central differences, h = 1e-6: the
    256 trajectories (+-h in each of the 128 parameters; the value is the mean of the first
    pair) advance together in one batched launch per step, one per workgroup.  One "step"
    is one TDVP step of all 256 trajectories; a full gradient at T = 10 is 200 of them."""
    import ttns_lib_amd
    from ttns_lib_amd import workloads as W
    from ttns_lib_amd.capi import SIGMA, Batch, Integrator, sb_problem
    from ttns_lib_amd.grad import perturbed_models

    lib = ttns_lib_amd.load_library()
    inst = W.c4_instance()
    models = perturbed_models(inst.model, 1e-6, include_base=False)
    keep, igs = [], []
    for m in models:
        tree, sop = sb_problem(lib, m)
        ig = Integrator(lib, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-14)
        ig.set_sb_initial_state(m)
        keep.append((tree, sop))
        igs.append(ig)
    batch = Batch(lib, igs)
    warm = max(1, args.warmup)
    t0 = time.perf_counter()
    batch.step(warm)
    print(f"[bench] c4grad warmup {warm} step(s) {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    f0 = sum(ig.flops() for ig in igs)
    t0 = time.perf_counter()
    batch.step(args.steps)
    wall = time.perf_counter() - t0
    kernel_ms = batch.last_step_ms()
    flops = sum(ig.flops() for ig in igs) - f0
    vals = np.array([ig.expect(0, SIGMA["z"]).real for ig in igs])
    n = len(inst.model.wk)
    d = (vals[0::2] - vals[1::2]) / 2e-6
    achieved = flops / (kernel_ms / 1e3) / 1e12
    out = {
        "metric": METRIC,
        "value": len(igs) * args.steps / wall,
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Ohmic bath discretisation, random orthonormal unoccupied functions)",
        "config": {"workload": "c4grad: central-difference gradient, 64-mode Ohmic alpha 0.2 wc 5 T=0, boson 12, "
                               "bond 24, dt 0.05, 256 trajectories in one batch", "instances": len(igs),
                   "krylov_dim": inst.krylov_dim, "krylov_tol": 1e-14},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "kernel": "ttnd::tdvp_run_kernel",
                     "flops_timed": flops},
        "gradient": {"T": float(igs[0].time()), "sigma_z": float(0.5 * (vals[0] + vals[1])),
                     "max_abs_dg": float(np.abs(d[:n]).max()), "max_abs_dw": float(np.abs(d[n:]).max()),
                     "projected_s_for_T10": 200 * wall / args.steps},
        "plain_gpu": None,
        "cpu_baseline": None,
    }
    batch.destroy()
    print(json.dumps(out))


def bench_c4adj(args):
    """Gradient workload by the costate method: d<sigma_z(T)>/d(g_k, w_k) for the 64-mode Ohmic
    bath (alpha 0.2, wc 5, T = 0, boson dim 12, bond dim 24, dt 0.05) over T = steps * dt with a
    checkpoint every --checkpoint-every steps (ttn_sb_adjoint_run).  One "step" is one step of
    the forward trajectory and its share of the backward pass; a full gradient at T = 10 is 200
    of them.  The timed region is the whole call (forward with checkpoints, recomputation,
    quadrature launches, transition densities)."""
    import ttns_lib_amd
    from ttns_lib_amd import workloads as W
    from ttns_lib_amd.capi import SIGMA, Integrator, sb_problem

    lib = ttns_lib_amd.load_library()
    inst = W.c4_instance()
    tree, sop = sb_problem(lib, inst.model)
    ig = Integrator(lib, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-14)
    ig.set_sb_initial_state(inst.model)
    every = max(1, min(args.checkpoint_every, args.steps))
    t0 = time.perf_counter()
    value, dg, dw = ig.sb_adjoint(inst.model, args.steps, 0, SIGMA["z"], every=every, nquad=0)
    wall = time.perf_counter() - t0
    ig.destroy()
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "gradient-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Ohmic bath discretisation, random orthonormal unoccupied functions)",
        "config": {"workload": "c4adj: costate gradient, 64-mode Ohmic alpha 0.2 wc 5 T=0, boson 12, bond 24, "
                               "dt 0.05", "checkpoint_every": every, "krylov_dim": inst.krylov_dim,
                   "krylov_tol": 1e-14},
        "gradient": {"T": args.steps * inst.dt, "sigma_z": value, "seconds": wall,
                     "max_abs_dg": float(np.abs(dg).max()), "max_abs_dw": float(np.abs(dw).max()),
                     "projected_s_for_T10": 200 * wall / args.steps},
        "roofline": None,
        "plain_gpu": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def bench_copy(args):
    """Calibration for HBM traffic counters: a device-to-device copy of 1 GiB (read 1 GiB,
    write 1 GiB per repetition) by torch's elementwise copy kernel."""
    import torch

    n = 1 << 27  # 2^27 float64 = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"workload": "copy", "bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n,
                      "copies": args.steps, "GBps": 2 * 8 * n * args.steps / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed steps: the first (the first-step ladder) in a launch of its own, then the rest")
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--plain-sample", type=int, default=8, help="instances stepped alone for the plain GPU rate")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle time for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--checkpoint-every", type=int, default=20, help="c4adj: steps between checkpoints")
    ap.add_argument("--report", action="store_true",
                    help="print each instance's team size, flops and device phase times of the timed steps to stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the batched sweep "
              "shards across GPUs by instance, which is not part of this build", file=sys.stderr)
        raise SystemExit(2)
    warm = max(1, args.warmup)
    if args.workload == "c2":
        bench_c2(args)
        return
    if args.workload == "copy":
        bench_copy(args)
        return
    if args.workload == "c4grad":
        bench_c4grad(args)
        return
    if args.workload == "c4adj":
        bench_c4adj(args)
        return

    import ttns_lib_amd
    from ttns_lib_amd.capi import Batch

    lib = ttns_lib_amd.load_library()
    insts, desc = workload(args.workload, lib)
    print(f"[bench] workload {args.workload}: {len(insts)} instances", file=sys.stderr, flush=True)
    igs = [make(lib, i) for i in insts]
    batch = Batch(lib, igs)
    t0 = time.perf_counter()
    # the first step (the first-step ladder, six sub-steps) in a launch of its own, then the rest:
    # the team-size autotuner sizes the next launch's teams from the last launch's device times,
    # which should be those of regular steps
    batch.step(1)
    if warm > 1:
        batch.step(warm - 1)
    print(f"[bench] warmup {warm} step(s) {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    # the plain-path sample: the slowest, fastest and median instances by device time so far,
    # checkpointed here so that each can later take the first timed step alone
    t_warm = np.array([ig.profile()["total"] for ig in igs])
    order = [int(i) for i in np.argsort(-t_warm)]
    k = min(max(0, args.plain_sample), len(insts))
    n_slow, n_fast = (k + 2) // 3, k // 3
    n_mid = k - n_slow - n_fast
    lo = len(order) // 2 - n_mid // 2
    pick = sorted(set(order[:n_slow]) | set(order[len(order) - n_fast:]) | set(order[lo:lo + n_mid])) \
        if k > 0 and len(insts) > 1 else []
    ckpt = {i: igs[i].checkpoint() for i in pick}
    f0 = np.array([ig.flops() for ig in igs])
    p0 = [ig.profile() for ig in igs]
    s0 = [ig.node_stats() for ig in igs] if args.report else None
    t0 = time.perf_counter()
    batch.step(args.steps)
    wall = time.perf_counter() - t0
    kernel_ms = batch.last_step_ms()
    f1 = np.array([ig.flops() for ig in igs])
    p1 = [ig.profile() for ig in igs]
    flops = float(np.sum(f1 - f0))
    per_inst_flops_step = (f1 - f0) / args.steps
    t_batch = np.array([(b["total"] - a["total"]) / 1e6 / args.steps for a, b in zip(p0, p1)])  # s per step
    value = len(insts) * args.steps / wall
    achieved = flops / (kernel_ms / 1e3) / 1e12
    print(f"[bench] timed {args.steps} step(s): wall {wall:.3f}s kernel {kernel_ms / 1e3:.3f}s "
          f"{flops / 1e9:.1f} GFLOP", file=sys.stderr, flush=True)
    if args.report:
        report(insts, igs, batch.teams(), f1 - f0, p0, args.steps, s0)
    sz = [float(ig.trajectory()[-1, 0]) for ig in igs]
    norm_dev = max(abs(float(ig.trajectory()[-1, 1]) - 1.0) for ig in igs)

    # plain per-instance path: the sample, restored from its checkpoints, one step alone each
    plain = None
    if pick:
        alone = {}
        for i in pick:
            ig = make(lib, insts[i])
            ig.restore(ckpt[i])
            t1 = time.perf_counter()
            ig.step(1)
            alone[i] = time.perf_counter() - t1
            ig.destroy()
            print(f"[bench] instance {i} alone: {alone[i]:.3f}s (in the batch {t_batch[i]:.3f}s)", file=sys.stderr,
                  flush=True)
        ratio = float(np.median([alone[i] / t_batch[i] for i in pick]))
        est = sum(alone.values()) + sum(t_batch[j] * ratio for j in range(len(insts)) if j not in alone)
        plain = {"value": len(insts) / est, "unit": "instance-steps/s",
                 "sample": f"instances {pick} (slowest, fastest and median in the batch) restored from checkpoints "
                           f"and stepped alone: {', '.join(f'{alone[i]:.2f}' for i in pick)} s for the step that "
                           f"took them {', '.join(f'{t_batch[i]:.2f}' for i in pick)} s of device time in the "
                           f"batch; the other {len(insts) - len(pick)} instances' alone times estimated as their "
                           f"batch device time x the median ratio {ratio:.3f} (extrapolated)",
                 "seconds_per_step_all_alone": est,
                 "batched_over_plain": value / (len(insts) / est)}
    elif len(insts) == 1:
        plain = {"value": value, "unit": "instance-steps/s", "sample": "single instance: batched == plain"}

    cpu, dsz = (None, None) if args.no_cpu else cpu_baseline(lib, insts, per_inst_flops_step, args.cpu_budget)
    cpu_all = None
    if not args.no_cpu and len(insts) > 1:
        cpu_all = cpu_baseline_all_cores(insts, per_inst_flops_step, min(16, os.cpu_count() or 1, len(insts)))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (frozen bath discretisations, random orthonormal unoccupied functions)",
        "config": {"workload": f"{args.workload}: {desc}", "instances": len(insts), "krylov_dim": 12,
                   "krylov_tol": 1e-12},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "ttnd::tdvp_run_kernel", "flops_timed": flops},
        "plain_gpu": plain,
        "cpu_baseline": cpu,
        "check": {"sigma_z_last": sz[:4], "max_norm_deviation": norm_dev, "oracle_max_dsigma_z": dsz},
        "batch_device_seconds_per_step": {"max": float(t_batch.max()), "median": float(np.median(t_batch)),
                                          "sum": float(t_batch.sum())},
    }
    if cpu:
        out["vs_cpu"] = value / cpu["value"]
    if cpu_all:
        out["cpu_all_cores"] = cpu_all
        out["vs_cpu_all_cores"] = value / cpu_all["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
