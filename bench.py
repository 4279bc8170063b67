"""Benchmark of the BUG hot path on one MI355X.

Prints ONE JSON line (rank 0).  Workload: the XXZ chain (delta = 0.5) at L = 128 on a seeded
synthetic U(1) state with bond dimension up to D = 1024 (Gaussian charge-sector weights, sigma 4),
canonicalised and normalised on the device.

``--workload bug_step`` (default): one step = one ``cbe_bug_step`` (order 2, tau = -0.02 i,
maxdim = D, work_ceiling_rule = maxdim, default tolerances) on the device-resident state.  The
headline form is ``--precision mixed`` (FP32-MFMA block contractions, FP64 Krylov / QR / SVD);
the FP64 drop-in is stepped from the same state beside it, timed on its own, and the largest
|<S^z_j>| difference between the two runs after the timed steps is reported as the error of the
mixed form.  ``--precision fp64`` times the drop-in alone.
``--workload heff_sweep``: one step = a left-to-right pass doing the one-site effective-Hamiltonian
apply and the left environment push at every site (the contraction core of a half-sweep).

The dominant kernel (grouped complex block GEMM of the one-site apply) is timed separately with
HIP events on its own stream for the roofline; the CPU baseline times the oracle's one-site apply
on the same environments (downloaded) with numpy/OpenBLAS on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

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
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--L", type=int, default=128)
    p.add_argument("--D", type=int, default=1024)
    p.add_argument("--delta", type=float, default=0.5)
    p.add_argument("--order", type=int, default=2)
    p.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    p.add_argument("--half-sweeps", choices=["serial", "parallel"], default="parallel",
                   help="run the two half-sweeps one after the other or concurrently")
    p.add_argument("--no-fp64", action="store_true",
                   help="with --precision mixed: skip the FP64 comparison run")
    p.add_argument("--workload", choices=["bug_step", "heff_sweep"], default="bug_step")
    p.add_argument("--kernel-reps", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--serving", type=int, default=12,
                   help="requests of the seeded Poisson serving mix to replay (0: skip)")
    p.add_argument("--serving-rate", type=float, default=1.0, help="arrivals per second")
    p.add_argument("--slots", type=int, default=4, help="concurrent requests of the serving loop")
    p.add_argument("--heston", type=int, default=4,
                   help="Heston instances (2^10 x 2^10 grid) stepped through the serving loop "
                        "(0: skip)")
    p.add_argument("--heston-steps", type=int, default=1)
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


def main():
    args = parse()
    # the step's own progress lines (bonds, closure, root substeps) on stderr: long steps stay visible
    os.environ.setdefault("BUG_PROGRESS", "1")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    import torch

    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo
    from bondupdategalerkin_amd.mps import random_mps

    torch.cuda.set_device(0)
    prec = _lib.PREC_FP64 if args.precision == "fp64" else _lib.PREC_FP32
    ckind = "fp64" if args.precision == "fp64" else "fp32"
    L, D = args.L, args.D
    t0 = time.time()
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import tensor_ops as T

    psi = random_mps(L, D, seed=0)
    mpo = xxz_mpo(L, delta=args.delta)
    mid = L // 2
    # mixed-canonical about the centre and normalised: the environments below stay bounded
    B.canonical(psi, mid)
    T.scale_(1.0 / T.norm(psi[mid]), psi[mid])
    torch.cuda.synchronize()
    print(f"# state built in {time.time() - t0:.1f}s; bond dims max {max(psi.bond_dims())}",
          file=sys.stderr, flush=True)

    if args.autotune == "table":
        if os.path.exists(_lib.SCHEDULE_TABLE):
            _lib.autotune_load(_lib.SCHEDULE_TABLE)
            _lib.set_autotune(_lib.AUTOTUNE_USE)
        else:
            print(f"# no tuning table at {_lib.SCHEDULE_TABLE}: heuristic schedules",
                  file=sys.stderr, flush=True)

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
    # for the shape and form they were measured on
    traffic = (126611 * 1024 if (ckind == "fp32" and L == 128 and D == 1024 and args.delta == 0.5
                                 and args.autotune == "table") else None)
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "kernel": (f"grouped_zgemm {ckind} (one-site H_eff apply: two grouped-GEMM "
                           "stages, each followed by its split-k reduce when split)"),
                "schedules": [{"tile": list(TILES[c]), "splits": sp} for c, sp in schedules],
                "flops_per_apply": flops_apply, "apply_ms": round(apply_ms, 4),
                "timing": "HIP events on the launch stream around each eager apply (its 2-3 "
                          "launches and the gaps between them), mean of kernel_reps",
                "apply_ms_graph": round(apply_ms_graph, 4), "graph_equals_eager": graph_equal}
    if args.kernel_only:
        print(json.dumps({"kernel_only": True, "roofline": roofline}), flush=True)
        return

    # ---- instance packing: many small independent applies in one launch per stage ----
    packing = None
    if args.packing > 0:
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
        # the same instances' Krylov exponentials (12 Lanczos vectors each) in lockstep
        from bondupdategalerkin_amd import krylov as KR
        from bondupdategalerkin_amd.packed import lanczos_expv_packed

        v0s = [it[3] for it in small]
        lz_seq = time_it(lambda: [KR.lanczos_expv(
            lambda v, a=a, w=w, b=b: ev.apply_one_site(a, w, b, v, precision=prec), -0.05j, x,
            maxiter=12, substeps=1, conv_tol=0.0) for (a, w, b, x, _) in small], reps=3)
        lz_pk = time_it(lambda: lanczos_expv_packed(packed, -0.05j, v0s, maxiter=12), reps=3)
        packing = {"workload": (f"{args.packing} independent one-site H_eff applies (XXZ L=24, "
                                "D=64 states, centre site) on vectors held in one packed buffer: "
                                "one apply call after another vs one PackedApply call (one grouped "
                                "launch per stage, merged plan reused); and their Krylov "
                                "exponentials (12 Lanczos vectors, tau=-0.05i) one after another vs "
                                "in lockstep over packed vectors"),
                   "sequential_ms": round(seq_ms, 4), "packed_ms": round(bat_ms, 4),
                   "speedup": round(seq_ms / bat_ms, 2),
                   "lanczos_sequential_ms": round(lz_seq, 3), "lanczos_packed_ms": round(lz_pk, 3),
                   "lanczos_speedup": round(lz_seq / lz_pk, 2)}
        print(f"# instance packing: {packing}", file=sys.stderr, flush=True)

    # ---- the timed step ----
    statuses, extra = [], {}
    if args.workload == "bug_step":
        from bondupdategalerkin_amd import mps as M

        B.canonical(psi, L - 1)
        T.scale_(1.0 / T.norm(psi[L - 1]), psi[L - 1])
        B.canonical(psi, 0)
        torch.cuda.synchronize()
        twin = psi.copy() if (args.precision == "mixed" and not args.no_fp64) else None

        def make_step(state, form, record):
            def step():
                info = B.cbe_bug_step(state, mpo, -0.02j, maxdim=D, work_ceiling_rule="maxdim",
                                      precision=form, order=args.order,
                                      parallel=args.half_sweeps == "parallel")
                if record is not None:
                    record.append(B.budget_status(info.ledger))
                phases = " ".join(f"{k} {v:.2f}" for k, v in info.seconds.items() if k != "step")
                print(f"# {form} step: {info.seconds['step']:.2f}s {B.budget_status(info.ledger)} "
                      f"{info.ledger.unmet} max bond {max(state.bond_dims())} ({phases})",
                      file=sys.stderr, flush=True)
            return step

        step = make_step(psi, args.precision, statuses)
        if twin is not None:
            fp64_statuses = []
            step64 = make_step(twin, "fp64", fp64_statuses)
            tuning(True)
            for _ in range(args.warmup):
                step64()
            torch.cuda.synchronize()
            fp64_statuses.clear()
            tuning(False)
            t1 = time.perf_counter()
            for _ in range(args.steps):
                step64()
            torch.cuda.synchronize()
            ms64 = 1e3 * (time.perf_counter() - t1) / args.steps
            extra["fp64"] = {"steps_per_s": round(1e3 / ms64, 4), "ms_per_step": round(ms64, 3),
                             "budget_status": fp64_statuses}
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
    statuses.clear()
    tuning(False)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    if args.workload == "bug_step" and twin is not None:
        dsz = float(np.abs(M.magnetisation(psi) - M.magnetisation(twin)).max())
        extra["fp64"]["max_abs_dSz_mixed_vs_fp64"] = dsz
        extra["fp64"]["speedup_mixed_vs_fp64"] = round(extra["fp64"]["ms_per_step"] / ms_per_step, 3)

    if args.serving > 0 and args.workload == "bug_step":
        from bondupdategalerkin_amd import serving as S

        srv = {"mix": (f"{args.serving} requests, Poisson arrivals at {args.serving_rate}/s "
                       "(seed 0): 60% XX/XXZ chains L in {16,24,32}, maxdim in {16,32,64}, 2-6 "
                       "steps, order 2; 40% Heston QTT 2^5-2^7 per axis, maxdim 16-32, 2-4 steps; Heston "
                       "requests run FP64 in both forms (non-Hermitian, stiff)"),
               "slots": args.slots}
        for form in (["mixed", "fp64"] if args.precision == "mixed" else ["fp64"]):
            print(f"# serving mix ({form})", file=sys.stderr, flush=True)
            srv[form] = S.serve_mix(S.poisson_mix(args.serving, args.serving_rate, seed=0),
                                    args.slots, form)
            print(f"# serving {form}: {srv[form]}", file=sys.stderr, flush=True)
        extra["serving"] = srv

    if args.heston > 0 and args.workload == "bug_step":
        from bondupdategalerkin_amd import serving as S

        print(f"# heston batch: {args.heston} instances", file=sys.stderr, flush=True)
        hrng = np.random.default_rng(0)
        eng = S.Engine(slots=args.slots)
        try:
            t0 = time.perf_counter()
            rids = []
            for _ in range(args.heston):
                prm = (hrng.uniform(0, 0.05), hrng.uniform(0, 0.05), hrng.uniform(0.5, 4),
                       hrng.uniform(0.02, 0.15), hrng.uniform(0.1, 0.8), hrng.uniform(-0.9, 0))
                rids.append(eng.submit(S.Request(
                    "heston", {"bits": 10, "params": prm, "maxdim": 128,
                               "steps": args.heston_steps, "dt": 0.02}, precision="fp64")))
            res = eng.wait(rids, timeout=1200)
            span = time.perf_counter() - t0
        finally:
            eng.close()
        lat = np.array([r.latency for r in res])
        extra["heston_batch"] = {
            "workload": (f"{args.heston} Heston QTT instances, 2^10 x 2^10 grid (x in [-1,1], v in "
                         "[0,0.5]), params drawn as the batched config (seed 0), psi0 = "
                         f"exp(0.5x) exp(-v), order 2, maxdim 128, tau 0.02, hermitian=false, "
                         f"{args.heston_steps} step(s) each, FP64, {args.slots} concurrent slots"),
            "instance_steps_per_s": round(sum(r.steps for r in res) / span, 4),
            "makespan_s": round(span, 2),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 2),
            "errors": [r.error for r in res if r.error],
            "budget_status": [s for r in res for s in r.statuses],
        }
        print(f"# heston batch: {extra['heston_batch']}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sample_sites, mpo, lenvs, renvs, mid, args)

    out = {
        "metric": METRIC,
        "value": round(1e3 / ms_per_step, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.precision == "fp64" else "c64 contractions / c128 Krylov, QR, SVD",
        "data": "synthetic",
        "config": {"workload": (f"cbe_bug_step order {args.order}, tau=-0.02i, maxdim=D, "
                                "work_ceiling_rule=maxdim, synthetic canonical U(1) state")
                   if args.workload == "bug_step" else
                   "heff_sweep: one-site H_eff apply + left env push at every site",
                   "L": L, "D": D, "model": f"xxz delta={args.delta}", "symmetry": "U1",
                   "precision": args.precision, "half_sweeps": args.half_sweeps,
                   "parallelism": "single GPU",
                   "budget_status": statuses},
        **extra,
        "autotune": {"mode": {"table": "package tuning table for gfx950 "
                                        "(bondupdategalerkin_amd/schedules_gfx950.txt)",
                              "tune": "tuned in the kernel warm-up and warm-up steps, table "
                                      "used in the timed steps",
                              "off": "heuristic schedules"}[args.autotune],
                     "classes": _lib.autotune_count()[0]},
        "instance_packing": packing,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    st = ctx.stats()
    out["plan_cache"] = {"plans": st["plans"], "hits": st["hits"], "misses": st["misses"],
                         "context": "main thread"}
    if args.save_schedules:
        _lib.autotune_save(args.save_schedules)
    print(json.dumps(out), flush=True)


def cpu_baseline(sites, mpo, lenvs, renvs, mid, args):
    """The oracle on the host cores, on a bounded sample of the same step, scaled to steps/s.

    Sample: whole bond visits of the left half-sweep at the centre bond (unique-QR frame, DD-QR CBE
    with its probe, streamed sketch and rank bisection, local Lanczos enrichment) on the device
    tensors downloaded, repeated for about ``--cpu-seconds``; plus one truncation cut (SVD of the
    centre site).  Estimate: an order-2 step visits every bond twice (half-sweeps, frontier sweep)
    and cuts every site once, so t_step ~ 2 (L-1) t_visit + L t_cut (an upper estimate: bonds near
    the ends are smaller than the centre one)."""
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
    A, B = fam(sites[0]), fam(sites[1])
    opts = obug.StepOptions(order=2, maxdim=D, work_ceiling_rule="maxdim")
    alloc = obug.basis_allocation(opts.tol_basis, L - 1, tau, 1.0, 2)
    wcap = 2 * D
    rng = ola.ProbeRNG(0x5EED)

    def visit():
        ledger = obug.Ledger()
        fr = obug.frame_from(A, B)
        ex = obug.dd_cbe_expand(fr, W, mid, Lf, Rf, "left", alloc, ledger, wcap, opts, rng)
        Kex = ola.site_times_core(A, ola.overlap_right(B, ex.V_ex))
        rch1 = ob.push_right(Rf, ex.V_ex, W[mid + 1])
        obug.krylov_enrich(Kex, lambda v: ob.apply_one_site(Lf, W[mid], rch1, v), ex.U_ex, "left",
                           tau, alloc, wcap, "lanczos")

    n, t0 = 0, time.perf_counter()
    while True:
        visit()
        n += 1
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    t_visit = (time.perf_counter() - t0) / n
    t1 = time.perf_counter()
    ola.svd_left(A, cutoff=0.0, nkeep=D)
    t_cut = time.perf_counter() - t1
    t_step = 2 * (L - 1) * t_visit + L * t_cut
    try:
        from threadpoolctl import threadpool_info
        threads = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:
        threads = os.cpu_count()
    return {"value": round(1.0 / t_step, 6), "unit": "steps/s (estimated from a sample)",
            "cores": threads, "kind": "port",
            "sample": (f"{n} oracle bond visits at the centre bond (frame, DD-QR CBE, Lanczos "
                       f"enrichment) of the L={L}, D={D} state, {1e3 * t_visit:.1f} ms each, and "
                       f"one centre truncation SVD, {1e3 * t_cut:.1f} ms; step ~ 2 (L-1) t_visit "
                       f"+ L t_cut = {t_step:.2f} s")}


if __name__ == "__main__":
    main()
