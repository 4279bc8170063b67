"""Benchmark of the BUG hot path on one MI355X.

Prints ONE JSON line.  Workload: the XXZ chain (delta = 0.5) at L = 128 on a seeded synthetic U(1)
state with bond dimension up to D = 1024 (Gaussian charge-sector weights, sigma 4; right-bond
columns weighted so that the Schmidt values fall by ``--tail`` across each bond), canonicalised
and normalised on the device.

``--workload bug_step`` (default): one step = one ``cbe_bug_step`` (order 2, tau = -0.02 i,
maxdim = D, work_ceiling_rule = maxdim, default tolerances, parallel half-sweeps) on the
device-resident state.  The headline form is ``--precision mixed`` (FP32-MFMA block contractions,
FP64 Krylov / QR / SVD): ``--warmup`` untimed steps, then exactly ``--steps`` timed steps.  The
FP64 drop-in is then stepped from the same starting state for as many steps as the time budget
holds (at most ``--fp64-steps`` after one warm-up) and the largest |<S^z_j>| difference between
the two forms after the same number of steps is reported as the error of the mixed form.
``--workload heff_sweep``: one step = a left-to-right pass doing the one-site effective-Hamiltonian
apply and the left environment push at every site (the contraction core of a half-sweep).

Everything after the timed steps (FP64 twin, serving mix, Heston batch, CPU baseline) runs under
one wall-clock budget (``--budget`` seconds from the start of the process); a phase that does not
fit is cut or skipped and the line says so.  A last-resort alarm prints the line with what exists
if the budget is overrun anyway.

The dominant kernel (grouped complex block GEMM of the one-site apply) is timed separately with
HIP events on its own stream for the roofline; the CPU baseline times the oracle on bounded
samples of the same step's phases on the host cores (one thread and all threads).
"""
from __future__ import annotations

import argparse
import json
import os
import signal
import sys
import time

T_START = time.time()

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "BUG steps/s L=128 D=1024 mixed vs FP64 @1e-6 err; p50/p99 request latency"
# tile configurations of the grouped GEMM (waves in m, waves in n, 16 x 16 MFMA tiles per wave)
TILES = [(2, 2, 2, 2), (1, 1, 2, 2), (2, 2, 1, 1), (2, 2, 4, 2), (1, 1, 1, 1)]
PEAK = {"fp64": 78.6, "fp32": 157.3}   # dense MFMA TFLOP/s, MI355X datasheet
HBM_PEAK = 8000.0                      # GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--L", type=int, default=128)
    p.add_argument("--D", type=int, default=1024)
    p.add_argument("--delta", type=float, default=0.5)
    p.add_argument("--order", type=int, default=2)
    p.add_argument("--tail", type=float, default=1e-9,
                   help="Schmidt-value fall across each bond of the synthetic state (0: flat)")
    p.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    p.add_argument("--half-sweeps", choices=["serial", "parallel"], default="parallel",
                   help="run the two half-sweeps one after the other or concurrently")
    p.add_argument("--budget", type=float, default=500.0,
                   help="wall-clock seconds from process start for the whole run; phases after "
                        "the timed steps are cut to fit")
    p.add_argument("--fp64-steps", type=int, default=1,
                   help="FP64 comparison steps after one warm-up (0: skip the FP64 twin)")
    p.add_argument("--workload", choices=["bug_step", "heff_sweep", "heston_batch", "c3_quench"],
                   default="bug_step",
                   help="bug_step: the headline; heff_sweep: the contraction core of a "
                        "half-sweep; heston_batch: --heston Heston instances stepped together "
                        "for --steps batched steps; c3_quench: the XXZ L=128 Neel quench "
                        "(maxdim 1024) for --steps steps in each form")
    p.add_argument("--driver", choices=["c", "python"], default="c",
                   help="c: the step-level C entry point bug_step (C++ orchestration); python: "
                        "the package's Python orchestration of the same kernels")
    p.add_argument("--kernel-reps", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=8.0,
                   help="seconds of oracle work per CPU sample (two samples: 1 thread, all)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--serving", type=int, default=12,
                   help="requests of the seeded Poisson serving mix to replay (0: skip)")
    p.add_argument("--serving-rate", type=float, default=1.0, help="arrivals per second")
    p.add_argument("--serving-scale", choices=["small", "full"], default="small")
    p.add_argument("--slots", type=int, default=4, help="concurrent requests of the serving loop")
    p.add_argument("--heston", type=int, default=64,
                   help="Heston instances (2^10 x 2^10 grid) stepped together through "
                        "bug_step_batch (0: skip)")
    p.add_argument("--heston-steps", type=int, default=2, help="batched Heston steps")
    p.add_argument("--heston-root", default="exp", choices=["exp", "trapezoid", "radau2", "radau3"],
                   help="root integrator of --workload heston_batch (a rational one runs unshifted)")
    p.add_argument("--heston-workers", type=int, default=16,
                   help="worker threads of the batched step's per-instance stages")
    p.add_argument("--autotune", choices=["table", "tune", "off"], default="table",
                   help="table: grouped-GEMM schedules from the package's tuning table for "
                        "gfx950 (heuristic for classes it lacks); tune: the autotuner times "
                        "candidate schedules during the kernel warm-up and the warm-up steps "
                        "and the timed steps use its table; off: heuristic schedules")
    p.add_argument("--save-schedules", default=None,
                   help="write the autotuner's table (plan class -> schedule) to this file")
    p.add_argument("--packing", type=int, default=16,
                   help="independent small applies timed one by one and packed (0: skip)")
    p.add_argument("--kernel-only", action="store_true",
                   help="run only the dominant-kernel timing (for profiling)")
    return p.parse_args()


def left() -> float:
    """Seconds left of the run's budget."""
    return ARGS.budget - (time.time() - T_START)


def note(msg: str):
    print(f"# {time.time() - T_START:7.1f}s {msg}", file=sys.stderr, flush=True)


OUT: dict = {}
_PRINTED = [False]


def emit():
    """Print the JSON line once (from the normal end or from the budget alarm)."""
    if _PRINTED[0]:
        return
    _PRINTED[0] = True
    print(json.dumps(OUT), flush=True)


def _alarm(signum, frame):
    OUT.setdefault("cut", []).append("budget alarm: the line was printed before the run ended")
    emit()
    os._exit(0 if "value" in OUT else 3)


def main():
    global ARGS
    args = ARGS = parse()
    # the step's own progress lines (bonds, closure, root substeps) on stderr: long steps stay visible
    os.environ.setdefault("BUG_PROGRESS", "1")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    signal.signal(signal.SIGALRM, _alarm)
    signal.alarm(max(int(args.budget + 70 - (time.time() - T_START)), 1))
    import torch

    if args.workload in ("heston_batch", "c3_quench"):
        torch.cuda.set_device(0)
        (heston_batch_run if args.workload == "heston_batch" else c3_quench_run)(args)
        signal.alarm(0)
        emit()
        return

    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo
    from bondupdategalerkin_amd.mps import random_mps

    torch.cuda.set_device(0)
    note("torch and the library loaded")
    prec = _lib.PREC_FP64 if args.precision == "fp64" else _lib.PREC_FP32
    ckind = "fp64" if args.precision == "fp64" else "fp32"
    L, D = args.L, args.D
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import tensor_ops as T

    OUT.update({
        "metric": METRIC, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128" if args.precision == "fp64" else "c64 contractions / c128 Krylov, QR, SVD",
        "data": "synthetic",
    })
    t0 = time.time()
    psi = random_mps(L, D, seed=0, tail=args.tail if args.tail > 0 else None)
    mpo = xxz_mpo(L, delta=args.delta)
    mid = L // 2
    # mixed-canonical about the centre and normalised: the environments below stay bounded
    B.canonical(psi, mid)
    T.scale_(1.0 / T.norm(psi[mid]), psi[mid])
    torch.cuda.synchronize()
    note(f"state built in {time.time() - t0:.1f}s; bond dims max {max(psi.bond_dims())}")

    if args.autotune == "table":
        if os.path.exists(_lib.SCHEDULE_TABLE):
            _lib.autotune_load(_lib.SCHEDULE_TABLE)
            _lib.set_autotune(_lib.AUTOTUNE_USE)
        else:
            note(f"no tuning table at {_lib.SCHEDULE_TABLE}: heuristic schedules")

    def tuning(on):
        if args.autotune == "tune":
            _lib.set_autotune(_lib.AUTOTUNE_TUNE if on else _lib.AUTOTUNE_USE)

    tuning(True)
    # right environments once (inputs of every apply)
    renvs = ev.right_env_stack(psi, mpo, downto=mid + 1, precision=prec)
    lenvs = ev.left_env_stack(psi, mpo, upto=mid, precision=prec)
    torch.cuda.synchronize()

    ctx = _lib.context()

    # ---- dominant kernel: one-site H_eff apply at the centre site ----
    Lm, Rm, X = lenvs[mid], renvs[mid + 1], psi[mid]
    sample_sites = (psi[mid], psi[mid + 1])      # the tensors the environments were built from
    ev.apply_one_site(Lm, mpo[mid], Rm, X, precision=prec)
    flops_apply = ctx.last_flops()
    schedules = ctx.last_schedules()
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.kernel_reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.kernel_reps)]
    for r in range(args.kernel_reps):
        starts[r].record(stream)
        ev.apply_one_site(Lm, mpo[mid], Rm, X, precision=prec)
        ends[r].record(stream)
    torch.cuda.synchronize()
    apply_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    # the same apply captured as a hipGraph and replayed (reported beside the eager time)
    from bondupdategalerkin_amd.graphs import CapturedApply

    cap = CapturedApply(lambda x: ev.apply_one_site(Lm, mpo[mid], Rm, x, precision=prec), X)
    y_eager = ev.apply_one_site(Lm, mpo[mid], Rm, X, precision=prec)
    graph_equal = bool(torch.equal(cap.replay().data, y_eager.data))
    for r in range(args.kernel_reps):
        starts[r].record(stream)
        cap.replay()
        ends[r].record(stream)
    torch.cuda.synchronize()
    cap.close()
    apply_ms_graph = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    achieved = flops_apply / (apply_ms * 1e-3) / 1e12
    peak = PEAK[ckind]
    # HBM bytes per apply, from the two PMC passes (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors
    # from a plain copy of known size) recorded in profiles/heff_apply_L128_D1024_mixed.md; only
    # for the shape and form they were measured on (the default state; 126611 KiB on the flat one)
    traffic = None
    if ckind == "fp32" and L == 128 and D == 1024 and args.delta == 0.5 and args.autotune == "table":
        traffic = {1e-9: 125643 * 1024, 0.0: 126611 * 1024}.get(args.tail)
    OUT["roofline"] = {
        "bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
        "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
        "kernel": (f"grouped_zgemm {ckind} (one-site H_eff apply at the centre site: two "
                   "grouped-GEMM stages, each followed by its split-k reduce when split)"),
        "schedules": [{"tile": list(TILES[c]), "splits": sp} for c, sp in schedules],
        "flops_per_apply": flops_apply, "apply_ms": round(apply_ms, 4),
        "timing": "HIP events on the launch stream around each eager apply (its 2-3 "
                  "launches and the gaps between them), mean of kernel_reps",
        "apply_ms_graph": round(apply_ms_graph, 4), "graph_equals_eager": graph_equal}
    note(f"roofline: {OUT['roofline']['achieved']} TFLOP/s ({apply_ms:.4f} ms per apply)")
    # the step's largest kernel family by time: the batched block SVD of a centre cut
    OUT["roofline_svd"] = svd_roofline(psi[mid], args.kernel_reps, stream)
    note(f"svd roofline: {OUT['roofline_svd']['achieved']} TFLOP/s "
         f"({OUT['roofline_svd']['svd_ms']:.2f} ms per factorization)")
    if args.kernel_only:
        OUT["kernel_only"] = True
        emit()
        return

    # ---- instance packing: many small independent applies in one launch per stage ----
    if args.packing > 0:
        OUT["instance_packing"] = packing_phase(args, prec, stream)

    # ---- the timed step ----
    statuses = []
    OUT["config"] = {
        "workload": (f"cbe_bug_step order {args.order}, tau=-0.02i, maxdim=D, "
                     "work_ceiling_rule=maxdim, default tolerances, synthetic canonical U(1) "
                     f"state (Schmidt values falling by {args.tail:g} across each bond)")
        if args.workload == "bug_step" else
        "heff_sweep: one-site H_eff apply + left env push at every site",
        "L": L, "D": D, "model": f"xxz delta={args.delta}", "symmetry": "U1",
        "precision": args.precision, "half_sweeps": args.half_sweeps,
        "parallelism": "single GPU", "budget_status": statuses}
    snap, twin = None, None
    ncmp = min(1 + args.fp64_steps, args.warmup + args.steps)
    if args.workload == "bug_step":
        from bondupdategalerkin_amd import mps as M

        B.canonical(psi, L - 1)
        T.scale_(1.0 / T.norm(psi[L - 1]), psi[L - 1])
        B.canonical(psi, 0)
        torch.cuda.synchronize()
        done = [0]
        phases_log = []
        kw = dict(maxdim=D, work_ceiling_rule="maxdim", precision=args.precision,
                  order=args.order, parallel=args.half_sweeps == "parallel")
        if args.driver == "c":
            from bondupdategalerkin_amd import cstep as C

            cstate, cmpo = C.CState(psi), C.CMPO(mpo)
            if args.precision == "mixed" and args.fp64_steps > 0:
                twin = cstate.clone()
            OUT["config"]["driver"] = ("bug_step, the step-level C entry point of libbug.so "
                                       "(the whole step orchestrated in C++)")

            def step():
                nonlocal snap
                info = C.bug_step(cstate, cmpo, -0.02j, **kw)
                done[0] += 1
                statuses.append(info["status"])
                phases_log.append({k: round(v, 3) for k, v in info["seconds"].items()})
                note(f"{args.precision} step {done[0]}: {info['seconds']['step']:.2f}s "
                     f"{statuses[-1]} {info['unmet_names']} max bond {info['r_final']} "
                     f"({' '.join(f'{k} {v:.2f}' for k, v in info['seconds'].items() if k != 'step')})")
                if twin is not None and done[0] == ncmp:
                    snap = cstate.clone()
        else:
            if args.precision == "mixed" and args.fp64_steps > 0:
                twin = psi.copy()
            OUT["config"]["driver"] = "cbe_bug_step, the package's Python orchestration"

            def step():
                nonlocal snap
                info = B.cbe_bug_step(psi, mpo, -0.02j, **kw)
                done[0] += 1
                statuses.append(B.budget_status(info.ledger))
                phases_log.append({k: round(v, 3) for k, v in info.seconds.items()})
                note(f"{args.precision} step {done[0]}: {info.seconds['step']:.2f}s "
                     f"{statuses[-1]} {info.ledger.unmet} max bond {max(psi.bond_dims())} "
                     f"({' '.join(f'{k} {v:.2f}' for k, v in info.seconds.items() if k != 'step')})")
                if twin is not None and done[0] == ncmp:
                    snap = psi.copy()
    else:
        def step():
            lenv = lenvs[0]
            for i in range(L):
                R = renvs[i + 1] if i + 1 < len(renvs) and renvs[i + 1] is not None else None
                if R is not None:
                    ev.apply_one_site(lenv, mpo[i], R, psi[i], precision=prec)
                lenv = ev.push_left(lenv, psi[i], mpo[i], precision=prec)

    tuning(True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    tuning(False)
    warm_statuses = list(statuses)
    statuses.clear()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    OUT["value"] = round(1e3 / ms_per_step, 4)
    OUT["ms_per_step"] = round(ms_per_step, 3)
    if args.workload == "bug_step":
        OUT["config"]["warmup_budget_status"] = warm_statuses
        OUT["config"]["max_bond_after"] = max(cstate.bond_dims() if args.driver == "c"
                                              else psi.bond_dims())
        OUT["phase_seconds_per_step"] = {
            k: round(float(np.mean([p.get(k, 0.0) for p in phases_log[args.warmup:]])), 3)
            for k in phases_log[-1]} if phases_log[args.warmup:] else None
    note(f"headline: {OUT['value']} steps/s ({ms_per_step:.1f} ms per step)")

    # ---- FP64 twin: the drop-in form from the same start, as far as the budget holds ----
    if twin is not None:
        if args.driver == "c":
            CMPO_HOLD[0] = cmpo
        # the mixed form's own step times at the step indices the FP64 twin times (2 .. ncmp)
        same = [p["step"] for p in phases_log[1:ncmp] if "step" in p]
        ms_same = 1e3 * float(np.mean(same)) if same else ms_per_step
        OUT["fp64"] = fp64_phase(args, B, M, mpo, twin, snap, ncmp, ms_same)

    if not args.no_cpu:
        if left() > 2.5 * args.cpu_seconds + 15:
            OUT["cpu_baseline"] = cpu_baseline(sample_sites, mpo, lenvs, renvs, mid, args)
        else:
            OUT["cpu_baseline"] = None
            OUT.setdefault("cut", []).append("cpu_baseline: no budget left")

    if args.serving > 0 and args.workload == "bug_step":
        OUT["serving"] = serving_phase(args)

    if args.heston > 0 and args.workload == "bug_step":
        OUT["heston_batch"] = heston_phase(args)

    OUT["autotune"] = {"mode": {"table": "package tuning table for gfx950 "
                                         "(bondupdategalerkin_amd/schedules_gfx950.txt)",
                                "tune": "tuned in the kernel warm-up and warm-up steps, table "
                                        "used in the timed steps",
                                "off": "heuristic schedules"}[args.autotune],
                       "classes": _lib.autotune_count()[0]}
    st = ctx.stats()
    OUT["plan_cache"] = {"plans": st["plans"], "hits": st["hits"], "misses": st["misses"],
                         "context": "main thread"}
    OUT["run_seconds"] = round(time.time() - T_START, 1)
    if args.save_schedules:
        _lib.autotune_save(args.save_schedules)
    signal.alarm(0)
    emit()


def svd_roofline(A, reps, stream):
    """The batched block SVD of one truncation cut (the centre site's left matricisation, every
    charge sector in one call: QR preconditioning, block-Jacobi sweeps, U = Q V), timed with HIP
    events on its launch stream.  Algorithmic flops per sector, M x N after transposition to
    M >= N: 16 (2 M N^2 - 2 N^3 / 3) for the Householder QR (R and Q), 44 N^3 per one-sided Jacobi
    sweep (N (N - 1) / 2 rotations, each three length-N inner products and the update of two
    columns of X and of the accumulated V) times the sector's recorded sweep count, and 8 M N^2
    for U = Q V."""
    import torch

    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import tensor_ops as T

    ctx = _lib.context()
    ctx.check(ctx._lib.bug_set_svd_variant(ctx.handle, 0, 1))
    fac = T.factor_left(A)
    torch.cuda.synchronize()
    mats = list(fac.mats)
    sweeps = np.zeros(max(len(mats), 1), np.int32)
    ctx.check(ctx._lib.bug_svd_sweeps(ctx.handle, len(mats), _lib.ptr32(sweeps)))
    ctx.check(ctx._lib.bug_set_svd_variant(ctx.handle, 0, 0))
    flops = 0.0
    for (m, n), sw in zip(mats, sweeps):
        M, N = max(m, n), min(m, n)
        flops += 16.0 * (2.0 * M * N * N - 2.0 * N ** 3 / 3) + 44.0 * N ** 3 * int(sw) \
            + 8.0 * M * N * N
    reps = max(2, min(reps, 5))
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    for r in range(reps):
        starts[r].record(stream)
        T.factor_left(A)
        ends[r].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    achieved = flops / (ms * 1e-3) / 1e12
    peak = PEAK["fp64"]
    widths = sorted((min(m, n) for m, n in mats), reverse=True)
    return {
        "bound": "fp64 (Jacobi rotations on the vector ALUs, QR reflectors and U = Q V on MFMA)",
        "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4), "traffic": None,
        "kernel": ("batched block SVD of the centre site's left matricisation: QR panels and "
                   "reflectors, Gram block-Jacobi rounds (gbj_gram_kernel + gbj_rotate_kernel) "
                   "for sectors >= 128 wide, svd_finish, U = Q V, phase"),
        "sectors": len(mats), "widths_top": widths[:6], "sweeps_max": int(sweeps.max()),
        "flops_per_svd": flops, "svd_ms": round(ms, 3),
        "timing": "HIP events on the launch stream around each factorization, mean of reps"}


def packing_phase(args, prec, stream):
    import torch

    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd import krylov as KR
    from bondupdategalerkin_amd.mpo import xxz_mpo
    from bondupdategalerkin_amd.mps import random_mps
    from bondupdategalerkin_amd.packed import lanczos_expv_packed

    small = []
    for s in range(args.packing):
        st = random_mps(24, 64, seed=100 + s)
        B.canonical(st, 12)
        ms = xxz_mpo(24, delta=args.delta)
        le = ev.left_env_stack(st, ms, upto=12, precision=prec)
        re = ev.right_env_stack(st, ms, downto=13, precision=prec)
        small.append((le[12], ms[12], re[13], st[12], 0.0))
    torch.cuda.synchronize()

    def time_it(fn, reps=10):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    packed = ev.PackedApply([it[:3] for it in small], [it[3].layout for it in small],
                            precision=prec)
    for dst, it in zip(packed.x, small):
        dst.data.copy_(it[3].data)
    seq_ms = time_it(lambda: [ev.apply_one_site(a, w, b, x, precision=prec)
                              for (a, w, b, _, _), x in zip(small, packed.x)])
    bat_ms = time_it(packed)
    v0s = [it[3] for it in small]
    lz_seq = time_it(lambda: [KR.lanczos_expv(
        lambda v, a=a, w=w, b=b: ev.apply_one_site(a, w, b, v, precision=prec), -0.05j, x,
        maxiter=12, substeps=1, conv_tol=0.0) for (a, w, b, x, _) in small], reps=3)
    lz_pk = time_it(lambda: lanczos_expv_packed(packed, -0.05j, v0s, maxiter=12), reps=3)
    out = {"workload": (f"{args.packing} independent one-site H_eff applies (XXZ L=24, "
                        "D=64 states, centre site) on vectors held in one packed buffer: "
                        "one apply call after another vs one PackedApply call (one grouped "
                        "launch per stage, merged plan reused); and their Krylov "
                        "exponentials (12 Lanczos vectors, tau=-0.05i) one after another vs "
                        "in lockstep over packed vectors"),
           "sequential_ms": round(seq_ms, 4), "packed_ms": round(bat_ms, 4),
           "speedup": round(seq_ms / bat_ms, 2),
           "lanczos_sequential_ms": round(lz_seq, 3), "lanczos_packed_ms": round(lz_pk, 3),
           "lanczos_speedup": round(lz_seq / lz_pk, 2)}
    note(f"instance packing: {out}")
    return out


def fp64_phase(args, B, M, mpo, twin, snap, ncmp, ms_mixed):
    """FP64 steps from the mixed run's starting state: one warm-up, then up to --fp64-steps timed
    ones while the budget holds (an FP64 step is taken to cost 1.3x a mixed one when planning).
    ``ms_mixed`` is the mixed form's mean step time over the same step indices (2 .. ncmp), so
    the speedup compares like with like: later steps of the synthetic state cost more than early
    ones (their closures run longer), whatever the precision."""
    import torch

    out = {"warmup": 1, "steps": 0}
    est = 1.3 * ms_mixed / 1e3
    if left() < 2 * est + 60:
        OUT.setdefault("cut", []).append("fp64 twin: no budget left")
        return {"skipped": "budget"}
    sts = []
    kw = dict(maxdim=ARGS.D, work_ceiling_rule="maxdim", precision="fp64", order=ARGS.order,
              parallel=ARGS.half_sweeps == "parallel")

    def step():
        """One FP64 step; returns its time as the step reports it (as the mixed steps' are)."""
        if ARGS.driver == "c":
            from bondupdategalerkin_amd import cstep as C

            info = C.bug_step(twin, CMPO_HOLD[0], -0.02j, **kw)
            sts.append(info["status"])
            note(f"fp64 step: {info['seconds']['step']:.2f}s {sts[-1]} {info['unmet_names']}")
            return info["seconds"]["step"]
        info = B.cbe_bug_step(twin, mpo, -0.02j, **kw)
        sts.append(B.budget_status(info.ledger))
        note(f"fp64 step: {info.seconds['step']:.2f}s {sts[-1]} {info.ledger.unmet}")
        return info.seconds["step"]

    step()
    torch.cuda.synchronize()
    times = []
    while out["steps"] < ncmp - 1 and left() > est + 60:
        times.append(step())
        torch.cuda.synchronize()
        out["steps"] += 1
    if times:
        ms64 = 1e3 * float(np.mean(times))
        out.update(steps_per_s=round(1e3 / ms64, 4), ms_per_step=round(ms64, 3),
                   mixed_ms_per_step_same_steps=round(ms_mixed, 3),
                   compared_steps=[2, 1 + len(times)],
                   speedup_mixed_vs_fp64=round(ms64 / ms_mixed, 3))
    out["budget_status"] = sts
    if snap is not None and 1 + out["steps"] == ncmp:
        if ARGS.driver == "c":
            dsz = float(np.abs(snap.expect_diagonal() - twin.expect_diagonal()).max())
        else:
            dsz = float(np.abs(M.magnetisation(snap) - M.magnetisation(twin)).max())
        out["max_abs_dSz_mixed_vs_fp64"] = dsz
        out["compared_after_steps"] = ncmp
    else:
        OUT.setdefault("cut", []).append(
            f"fp64 twin: {1 + out['steps']} of {ncmp} steps in budget, no S^z comparison")
    note(f"fp64: {out}")
    return out


def serving_phase(args):
    from bondupdategalerkin_amd import serving as S

    mixdesc = {"small": ("60% XX/XXZ chains L in {16,24,32}, maxdim in {16,32,64}, 2-6 steps, "
                         "order 2; 40% Heston QTT 2^5-2^7 per axis, maxdim 16-32, 2-4 steps"),
               "full": ("60% XX/XXZ chains L in {32,64,96,128}, maxdim in {64,256,512,1024}, "
                        "20-200 steps; 40% Heston QTT 2^8-2^11 per axis, maxdim 32-128, 10-50 "
                        "steps")}[args.serving_scale]
    srv = {"mix": (f"{args.serving} requests, Poisson arrivals at {args.serving_rate}/s (seed 0): "
                   f"{mixdesc}; Heston requests run FP64 in both forms (non-Hermitian, stiff)"),
           "slots": args.slots}
    forms = ["mixed", "fp64"] if ARGS.precision == "mixed" else ["fp64"]
    for i, form in enumerate(forms):
        # an equal share of what is left for each form, keeping 10 s for the later phases
        share = (left() - 10) / (len(forms) - i)
        if share < 20:
            srv[form] = {"skipped": "budget"}
            OUT.setdefault("cut", []).append(f"serving {form}: no budget left")
            continue
        note(f"serving mix ({form}), {share:.0f}s at most")
        srv[form] = S.serve_mix(S.poisson_mix(args.serving, args.serving_rate, seed=0,
                                              scale=args.serving_scale),
                                args.slots, form, timeout=share)
        if srv[form].get("cancelled"):
            OUT.setdefault("cut", []).append(
                f"serving {form}: {srv[form]['cancelled']} requests cancelled at the budget")
        note(f"serving {form}: {srv[form]}")
    return srv


def heston_phase(args):
    """The batched Heston form: ``--heston`` instances on the 2^10 x 2^10 grid (parameter sets
    drawn as the batched configuration describes, seed 0) advanced together through
    ``bug_step_batch`` (stages before the root per instance on worker threads, the root
    exponentials as one lockstep Krylov recurrence over grouped launches), ``--heston-steps``
    batched steps within the budget; and one instance stepped alone through ``bug_step`` for the
    sequential rate."""
    import torch

    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import qtt as Q

    if left() < 60:
        OUT.setdefault("cut", []).append("heston batch: no budget left")
        return {"skipped": "budget"}
    note(f"heston batch: {args.heston} instances")
    hrng = np.random.default_rng(0)
    g = Q.qtt_grid(10, 10, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)
    init = Q.qtt_separable(g, lambda x: float(np.exp(0.5 * x)), lambda v: float(np.exp(-v)))
    prms = [(hrng.uniform(0, 0.05), hrng.uniform(0, 0.05), hrng.uniform(0.5, 4),
             hrng.uniform(0.02, 0.15), hrng.uniform(0.1, 0.8), hrng.uniform(-0.9, 0))
            for _ in range(args.heston)]
    kw = dict(order=2, maxdim=128, hermitian=False, work_ceiling_rule="maxdim",
              precision="fp64")
    if args.heston_root != "exp":
        kw.update(root_method=args.heston_root, root_shift=False, root_maxiter=200)
    states = [C.CState(Q.qtt_mps(init)) for _ in prms]
    mpos = [C.CMPO(Q.heston_mpo(g, Q.HestonParams(*p))) for p in prms]
    torch.cuda.synchronize()
    out = {"workload": (f"{args.heston} Heston QTT instances, 2^10 x 2^10 grid (x in [-1,1], v in "
                        "[0,0.5], non-periodic), params drawn as the batched config (seed 0), "
                        "psi0 = exp(0.5x) exp(-v), order 2, maxdim 128, tau 0.02, "
                        "hermitian=false, FP64, one bug_step_batch call per step "
                        f"({args.heston_workers} worker threads)"),
           "steps": 0, "seconds": [], "budget_status": [], "phase_seconds": []}
    active, dropped = list(range(len(states))), []
    out["dropped"] = dropped
    for _ in range(args.heston_steps):
        if out["seconds"] and left() < 1.5 * out["seconds"][-1] + 30:
            OUT.setdefault("cut", []).append(f"heston batch: {out['steps']} of "
                                             f"{args.heston_steps} batched steps in budget")
            break
        t0 = time.perf_counter()
        infos = batched_step(C, states, mpos, active, dropped, kw, args.heston_workers)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if not infos:
            break
        out["steps"] += 1
        out.setdefault("instances_per_step", []).append(len(infos))
        out["seconds"].append(round(dt, 2))
        out["budget_status"].append(sorted({i["status"] for i in infos}))
        out["phase_seconds"].append({"root_lockstep": round(infos[0]["seconds"]["root"], 2),
                                     "max_step": round(max(i["seconds"]["step"] for i in infos), 2)})
        note(f"heston batched step {out['steps']}: {dt:.1f}s, root lockstep "
             f"{infos[0]['seconds']['root']:.2f}s, expv iters max "
             f"{max(i['expv_iters'] for i in infos)}")
    if out["seconds"]:
        out["instance_steps_per_s"] = round(float(np.sum(out["instances_per_step"]))
                                            / float(np.sum(out["seconds"])), 4)
        out["max_bond_after"] = max(max(s.bond_dims()) for s in states)
    # one instance alone through bug_step, for the sequential rate
    if left() > 40:
        one = C.CState(Q.qtt_mps(init))
        t0 = time.perf_counter()
        info = C.bug_step(one, mpos[0], 0.02, **kw)
        torch.cuda.synchronize()
        out["single_instance_step_s"] = round(time.perf_counter() - t0, 2)
        out["single_instance_status"] = info["status"]
        if "instance_steps_per_s" in out:
            out["speedup_vs_one_at_a_time"] = round(
                out["instance_steps_per_s"] * out["single_instance_step_s"], 2)
    note(f"heston batch: {out}")
    return out


def batched_step(C, states, mpos, active, dropped, kw, workers):
    """One ``bug_step_batch`` over the active instances.  An instance whose step fails (a stiff
    parameter set whose state underflows to zero and then turns non-finite, as the oracle's own
    step does on it) is dropped from the batch and the batched step retried: a failed call changes
    no state.  Returns the reports of the instances stepped."""
    import re

    from bondupdategalerkin_amd import _lib

    while active:
        try:
            return C.bug_step_batch([states[i] for i in active], [mpos[i] for i in active], 0.02,
                                    workers=workers, **kw)
        except _lib.BugError as e:
            m = re.search(r"instance (\d+):", str(e))
            if m is None:
                raise
            bad = active[int(m.group(1))]
            dropped.append({"instance": bad, "error": re.sub(r"^.*instance \d+: ", "", str(e))})
            active.remove(bad)
            note(f"dropped instance {bad}: {e}")
    return []


def heston_batch_run(args):
    """The batched Heston configuration on its own: ``--heston`` instances (2^10 x 2^10 grid,
    parameter sets drawn as the configuration describes, seed 0), ``--warmup`` untimed and
    ``--steps`` timed ``bug_step_batch`` calls; value = instance-steps/s."""
    import torch

    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import qtt as Q

    hrng = np.random.default_rng(0)
    g = Q.qtt_grid(10, 10, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)
    init = Q.qtt_separable(g, lambda x: float(np.exp(0.5 * x)), lambda v: float(np.exp(-v)))
    prms = [(hrng.uniform(0, 0.05), hrng.uniform(0, 0.05), hrng.uniform(0.5, 4),
             hrng.uniform(0.02, 0.15), hrng.uniform(0.1, 0.8), hrng.uniform(-0.9, 0))
            for _ in range(args.heston)]
    kw = dict(order=2, maxdim=128, hermitian=False, work_ceiling_rule="maxdim",
              precision="fp64")
    states = [C.CState(Q.qtt_mps(init)) for _ in prms]
    mpos = [C.CMPO(Q.heston_mpo(g, Q.HestonParams(*p))) for p in prms]
    torch.cuda.synchronize()
    OUT.update({"metric": "Heston QTT instance-steps/s (batched)", "unit": "instance-steps/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "c128", "data": "synthetic",
                "config": {"workload": (f"{args.heston} Heston QTT instances, 2^10 x 2^10 grid, "
                                        "order 2, maxdim 128, tau 0.02, hermitian=false, one "
                                        "bug_step_batch call per step"),
                           "instances": args.heston, "workers": args.heston_workers,
                           "root": args.heston_root}})
    statuses, secs, counts, unmet = [], [], [], {}
    active, dropped = list(range(len(states))), []
    for k in range(args.warmup + args.steps):
        if left() < 30:
            OUT.setdefault("cut", []).append(f"heston batch: {len(secs)} of {args.steps} timed steps")
            break
        t0 = time.perf_counter()
        infos = batched_step(C, states, mpos, active, dropped, kw, args.heston_workers)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if not infos:
            break
        if k >= args.warmup:
            secs.append(dt)
            counts.append(len(infos))
            statuses.append(sum(i["status"] == "met" for i in infos))
            unmet = {}
            for i in infos:
                for name in i["unmet_names"]:
                    unmet[name] = unmet.get(name, 0) + 1
        note(f"batched step {k + 1}: {dt:.2f}s ({len(infos)} instances) root "
             f"{infos[0]['seconds']['root']:.2f}s met {statuses[-1] if statuses else 0}/"
             f"{len(infos)} max bond {max(i['r_final'] for i in infos)}")
    OUT["steps"] = len(secs)
    OUT["ms_per_step"] = round(1e3 * float(np.mean(secs)), 1) if secs else None
    OUT["value"] = round(float(np.sum(counts)) / float(np.sum(secs)), 4) if secs else None
    OUT["met_per_step"] = statuses
    OUT["instances_per_step"] = counts
    OUT["unmet_last_step"] = unmet
    OUT["dropped"] = dropped


def c3_quench_run(args):
    """The C3 configuration: XXZ (delta = 0.5) L = 128 Neel quench, maxdim 1024, work ceiling
    rule maxdim, dt 0.02, order 2, parallel half-sweeps, stepped ``--steps`` times in the mixed
    form and the FP64 drop-in side by side; reports the verdicts, the largest |dS^z| between the
    two forms along the way, the bond dimensions and the time of each form."""
    import torch

    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps as M
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 128
    mpo = C.CMPO(xxz_mpo(L, delta=args.delta))
    kw = dict(order=2, maxdim=1024, work_ceiling_rule="maxdim", parallel=True)
    sm, sf = C.CState(M.neel_state(L)), C.CState(M.neel_state(L))
    tm = tf = 0.0
    unmet, dsz_log, bonds = [], [], []
    done = 0
    for s in range(args.steps):
        if left() < 30:
            OUT.setdefault("cut", []).append(f"c3 quench: {done} of {args.steps} steps in budget")
            break
        t0 = time.perf_counter()
        im = C.bug_step(sm, mpo, -0.02j, precision="mixed", **kw)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        i6 = C.bug_step(sf, mpo, -0.02j, precision="fp64", **kw)
        torch.cuda.synchronize()
        tm += t1 - t0
        tf += time.perf_counter() - t1
        done += 1
        if im["status"] != "met" or i6["status"] != "met":
            unmet.append({"step": s + 1, "mixed": im["unmet_names"], "fp64": i6["unmet_names"]})
        dsz = float(np.abs(sm.expect_diagonal() - sf.expect_diagonal()).max())
        dsz_log.append(dsz)
        bonds.append((im["r_final"], i6["r_final"]))
        note(f"c3 step {s + 1}: mixed {t1 - t0:.2f}s {im['status']} fp64 "
             f"{time.perf_counter() - t1:.2f}s {i6['status']} bonds {bonds[-1]} dSz {dsz:.2e}")
    OUT.update({"metric": "BUG steps/s, C3 Neel quench (XXZ L=128, maxdim 1024)",
                "unit": "steps/s", "n_gpus": 1, "steps": done, "warmup": 0,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "c64 contractions / c128 Krylov, QR, SVD", "data": "Neel product state",
                "value": round(done / tm, 4) if tm > 0 else None,
                "fp64_steps_per_s": round(done / tf, 4) if tf > 0 else None,
                "config": {"workload": "XXZ delta=0.5 L=128 Neel quench, maxdim 1024, "
                                       "work_ceiling_rule maxdim, dt 0.02, order 2, parallel"},
                "unmet_steps": unmet, "first_unmet_step": unmet[0]["step"] if unmet else None,
                "max_abs_dSz_mixed_vs_fp64": max(dsz_log) if dsz_log else None,
                "dSz_every_10": [round(x, 10) for x in dsz_log[9::10]],
                "bonds_every_10": bonds[9::10],
                "seconds": {"mixed": round(tm, 1), "fp64": round(tf, 1)}})


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sites, mpo, lenvs, renvs, mid, args):
    """The oracle on the host cores, on bounded samples of the same step, scaled to steps/s.

    Samples (each repeated for about ``--cpu-seconds``, once with the BLAS pool held to one thread
    and once with every host thread): whole bond visits of the left half-sweep at the centre bond
    (unique-QR frame, DD-QR CBE with its probe, streamed sketch and rank bisection, local Lanczos
    enrichment) on the device tensors downloaded; and one truncation cut (SVD of the centre
    site).  Scale: an order-2 step visits every bond twice (half-sweeps, frontier sweep) and cuts
    every site once, so t_step ~ 2 (L-1) t_visit + L t_cut (an upper estimate: bonds near the
    ends are smaller than the centre one)."""
    from threadpoolctl import threadpool_info, threadpool_limits

    from oracle import blocks as ob
    from oracle import bug as obug
    from oracle import linalg as ola

    def fam(t):
        lay = t.layout
        return ob.Fam(lay.row.as_dict(), lay.col.as_dict(), lay.shifts, t.to_blocks())

    def site(W):
        return ob.MpoSite(W.cl, W.cr, zip(W.a, W.b, W.sout, W.sin, W.coeff))

    L, D = args.L, args.D
    tau = -0.02j
    W = {mid: site(mpo[mid]), mid + 1: site(mpo[mid + 1])}
    Lf, Rf = fam(lenvs[mid]), fam(renvs[mid + 2])
    A, Bs = fam(sites[0]), fam(sites[1])
    opts = obug.StepOptions(order=2, maxdim=D, work_ceiling_rule="maxdim")
    alloc = obug.basis_allocation(opts.tol_basis, L - 1, tau, 1.0, 2)
    wcap = 2 * D

    def visit(rng):
        ledger = obug.Ledger()
        fr = obug.frame_from(A, Bs)
        ex = obug.dd_cbe_expand(fr, W, mid, Lf, Rf, "left", alloc, ledger, wcap, opts, rng)
        Kex = ola.site_times_core(A, ola.overlap_right(Bs, ex.V_ex))
        rch1 = ob.push_right(Rf, ex.V_ex, W[mid + 1])
        obug.krylov_enrich(Kex, lambda v: ob.apply_one_site(Lf, W[mid], rch1, v), ex.U_ex, "left",
                           tau, alloc, wcap, "lanczos")

    def sample():
        rng = ola.ProbeRNG(0x5EED)
        n, t0 = 0, time.perf_counter()
        while True:
            visit(rng)
            n += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        t_visit = (time.perf_counter() - t0) / n
        t1 = time.perf_counter()
        ola.svd_left(A, cutoff=0.0, nkeep=D)
        t_cut = time.perf_counter() - t1
        return n, t_visit, t_cut

    threads = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    with threadpool_limits(limits=1):
        n1, v1, c1 = sample()
    nall, vall, call = sample()
    step1 = 2 * (L - 1) * v1 + L * c1
    stepall = 2 * (L - 1) * vall + L * call
    out = {"value": round(1.0 / stepall, 6), "unit": "steps/s (scaled from a sample)",
           "cores": threads, "kind": "port", "cpu_model": cpu_model(),
           "host_threads": os.cpu_count(),
           "one_thread": {"value": round(1.0 / step1, 6), "cores": 1,
                          "visit_ms": round(1e3 * v1, 1), "cut_ms": round(1e3 * c1, 1),
                          "visits": n1},
           "all_threads": {"value": round(1.0 / stepall, 6), "cores": threads,
                           "visit_ms": round(1e3 * vall, 1), "cut_ms": round(1e3 * call, 1),
                           "visits": nall},
           "sample": (f"oracle (numpy + OpenBLAS) bond visits at the centre bond (frame, DD-QR "
                      f"CBE, Lanczos enrichment) of the L={L}, D={D} state, repeated for "
                      f"{args.cpu_seconds:g} s, and one centre truncation SVD, on 1 thread and on "
                      f"{threads} threads; step ~ 2 (L-1) t_visit + L t_cut")}
    note(f"cpu baseline: {out}")
    return out


ARGS = None
CMPO_HOLD = [None]

if __name__ == "__main__":
    main()
