#!/usr/bin/env python3
"""Exposure-engine hot path on one MI355X: trade x path x date revaluations per second.

Workload (default, the largest configuration that fits one GPU): the batched uneven portfolio --
500 netting sets with Zipf(1.2) trade counts clipped to [1, 5,000] (20,744 trades: two-leg IRS in
3 currencies, FX forwards, 1,009 Bermudan swaptions), every other set under CSA, 3-currency
Hull-White + 2 lognormal FX + one CIR++ counterparty intensity, 200,000 paths x 120 dates, PFE
97.5%. The sets form one portfolio book (the reference's book semantics: every set on the book's
union grids, pathwise identical scenarios), compiled as waves of 50 netting sets (one plan each,
one shared workspace, the whole book's grids). A step is one execution of every wave: scenario
simulation, revaluation (American Monte Carlo for the Bermudans), collateral,
EE/ENE/EE*/ENE*/CVA increments and the exact PFE select, with the replayed normal stream and the
plan tables already resident in HBM.

Prints one JSON line. `roofline` is that of the dominant kernel (the path kernel), timed with
HIP events on the stream it runs on; `cpu_baseline` is the FP64 oracle (the reference engine's
algorithm, restated) timed on a bounded sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from zlatko_amd import abi, native  # noqa: E402
from zlatko_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_SIMDS = 1024          # 256 CUs x 4 SIMDs
VALU_CLOCK_HZ = 2.4e9      # peak engine clock
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
METRIC = "trade·path·date revals/sec; CVA wall time at ≤0.5% error, plain vs surrogate"
# Workload name -> (spec builder taking a path count, default path count).
WORKLOADS = {
    "c1": (lambda p: W.c1(num_paths=p), 5000),
    "c2": (lambda p: W.c2(num_paths=p), 50000),
    "c3": (lambda p: W.c3(num_paths=p), 100000),
    "c4": (lambda p: W.c4(num_paths=p), 100000),
    "c5": (lambda p: W.c5(num_paths=p), 200000),
}


WORKLOAD_TEXT = {
    "c1": "c1: 10 two-leg IRS (semiannual fixed, quarterly float), 1-factor Hull-White, 5,000 paths x 120 monthly dates",
    "c2": "c2: 3-ccy Hull-White + 2 FX, 1000 trades (700 IRS, 300 FX forwards), 20 CSA netting sets, "
          "50,000 paths x 260 weekly dates",
    "c3": "c3: 200 Bermudan swaptions (1-10y into 5-10y, annual exercise, degree-3 regressions), 1-factor "
          "Hull-White, 100,000 paths x 120 dates",
    "c4": "c4: 20 two-leg IRS + 10 FX forwards + 10 Bermudans, HW + FX + CIR++ counterparty / own intensities "
          "(0.3 wrong-way), 100,000 paths x 120 dates",
    "c5": "c5: 500 netting sets, Zipf(1.2) trade counts 1..5000 (20,744 trades: 3-ccy IRS, FX forwards, 1,009 "
          "Bermudans), half under CSA, 3-ccy HW + 2 FX + shared CIR++, 200,000 paths x 120 dates; one portfolio book "
          "(union grids, shared scenarios) in netting-set waves (one plan each, one shared workspace)",
}


def profile_figure(workload: str, precision: str, key: str):
    """A per-launch figure of the path kernel from the committed counter summary of this
    workload and precision (the line '<key>: N'), or None. traffic_bytes: HBM bytes (FETCH_SIZE
    and WRITE_SIZE passes of rocprofv3, each scaled by the copy-kernel calibration stated there);
    valu_instructions: SQ_INSTS_VALU."""
    f = ROOT / "profiles" / f"path_kernel_{workload}_{precision}.md"
    if not f.exists():
        return None, None
    for line in f.read_text().splitlines():
        if line.strip().startswith(key + ":"):
            return float(line.split(":", 1)[1].split()[0]), str(f.relative_to(ROOT))
    return None, None


def measured_traffic(workload: str, precision: str):
    return profile_figure(workload, precision, "traffic_bytes")


def describe(spec: abi.Spec, infos, waves: int) -> dict:
    d = {
        "workload": WORKLOAD_TEXT.get(spec.name, spec.name),
        "paths": int(infos[0].num_paths),
        "netting_sets": int(sum(i.num_sets for i in infos)),
        "trades": int(sum(len(s.trades) for s in spec.sets)),
        "max_simulation_dates": int(max(i.max_sim_dates for i in infos)),
        "max_reporting_dates": int(max(i.max_reporting_dates for i in infos)),
        "normals_per_step": int(infos[0].num_normals),
        "merged_terms": int(sum(i.total_terms for i in infos)),
        "pfe_quantile": spec.pfe_quantile,
    }
    if waves > 1:
        d["netting_set_waves"] = waves
    return d


def path_kernel_bytes(plans, outs, infos, elem: int) -> int:
    """Algorithmic HBM bytes of the path kernel per step: the replayed normal stream read once per
    plan (a book plan's sets replay one slice; independent sets each their own, counted once by
    the longest) plus max(V - C, 0) written once per (set, reporting date, path)."""
    stream = sum(i.stream_normals for i in infos) * elem
    rows = sum(o.c.num_dates for o in outs) * infos[0].num_paths * elem
    return stream + rows


def split_waves(spec: abi.Spec, wave_sets: int):
    """The netting sets in waves of wave_sets (one plan each, executed in order in one step),
    largest sets first so that the first wave's stream (its longest grid) serves the others."""
    if wave_sets <= 0 or wave_sets >= len(spec.sets):
        return [spec]
    order = sorted(spec.sets, key=lambda s: -len(s.trades))
    return [abi.Spec(spec.model, order[i:i + wave_sets], spec.base_grid, spec.num_paths, int(spec.config.seed),
                     bool(spec.config.antithetic), spec.pfe_quantile, spec.name)
            for i in range(0, len(order), wave_sets)]


def time_gpu(engine, plans, steps: int, warmup: int):
    # engine.synchronize() waits for the library's stream and then the whole device
    # (hipDeviceSynchronize): the timed region is bracketed by full-device synchronisation.
    for _ in range(warmup):
        for plan in plans:
            plan.execute_async()
    engine.synchronize()
    sums = {"path": [], "total": [], "amc": [], "reduce": [], "pfe": []}
    t0 = time.perf_counter()
    for _ in range(steps):
        step = dict.fromkeys(sums, 0.0)
        for plan in plans:
            plan.execute_async()
            t = engine.timing()  # waits for this wave's last kernel, reads its HIP events
            step["path"] += t.path_ms
            step["total"] += t.total_ms
            step["amc"] += t.amc_ms
            step["reduce"] += t.reduce_ms
            step["pfe"] += t.pfe_ms
        for k in sums:
            sums[k].append(step[k])
    engine.synchronize()
    wall = time.perf_counter() - t0
    return wall, {k: float(np.mean(v)) for k, v in sums.items()}, engine.timing()


def cva_wall_time_c4(engine, steps: int) -> dict:
    """The credit + Bermudan workload (c4, 100,000 paths): GPU time of one CVA -- one execute of the
    plan (simulation, American Monte Carlo or the surrogate, exposures, pathwise CVA increments) --
    on the plain FP64 path and on the FP32 path with the MLP surrogate valuing the Bermudans where
    the gate (hold-out EE error <= 0.5% of the peak EE) keeps it, with the CVA error against the
    plain one. The surrogate is trained and gated on two independent 20,000-path plans (training
    time reported apart)."""
    # The portfolio runs before leave the context in book mode with the book's grids and a shared
    # workspace; this workload is one stand-alone netting set.
    engine.set_book(False)
    engine.set_book_grids(None, None)
    engine.set_shared_workspace(False)
    tr = engine.compile(W.c4(num_paths=20000, seed=101))
    ho = engine.compile(W.c4(num_paths=20000, seed=202))
    t0 = time.perf_counter()
    sur = engine.surrogate_train(tr, ho, train_paths=2000, epochs=80, lr=5e-4, batch=4096, gate=0.005)
    train_s = time.perf_counter() - t0
    tr.close()
    ho.close()
    spec = W.c4()
    p64, p32 = engine.compile(spec, abi.CCR_FP64), engine.compile(spec, abi.CCR_FP32)

    def timed(run):
        run()
        engine.synchronize()
        ms = []
        for _ in range(steps):
            run()
            ms.append(engine.timing().total_ms)
        return float(np.mean(ms))

    plain_ms = timed(p64.execute_async)
    fp32_ms = timed(p32.execute_async)
    sur_ms = timed(lambda: p32.execute_surrogate_async(sur))
    ref = p64.execute()[0]
    out = p32.execute_surrogate(sur)[0]
    cva_ref = float(ref.get("cva_increments")[1:].sum())
    cva = float(out.get("cva_increments")[1:].sum())
    rep = sur.report
    for p in (p64, p32):
        p.close()
    sur.close()
    return {"workload": WORKLOAD_TEXT["c4"], "paths": int(spec.num_paths), "plain_fp64_ms": plain_ms,
            "fp32_regression_ms": fp32_ms, "fp32_surrogate_ms": sur_ms, "speedup_surrogate_vs_plain": plain_ms / sur_ms,
            "cva_plain_fp64": cva_ref, "cva_fp32_surrogate": cva, "cva_rel_err": abs(cva - cva_ref) / abs(cva_ref),
            "surrogate_sets": f"{rep.enabled_sets}/{rep.candidate_sets}", "gate_ee": 0.005,
            "holdout_ee_err": rep.max_ee_error, "train_s": train_s}


def cva_wall_time_c5(engine, specs, steps: int, outs64, ms64: float, ms32: float, gate: float = 0.005) -> dict:
    """The metric's second figure on the batched portfolio itself: GPU time of one CVA of the whole
    book (every wave executed once) on the plain FP64 path against the FP32 path with the MLP
    surrogate valuing the Bermudans. Each wave (a plan of 50 netting sets on the book's scenario)
    gets one network with an output per set holding Bermudans, trained and gated on two independent
    8,000-path plans of the same sets. A wave takes the network only where its Bermudan stage is
    faster than the regressions on the production plan, and -- for the gated figure -- only where
    every set of the wave passes the gate (hold-out EE error <= 0.5% of the set's peak EE; a set
    that fails keeps the regressions, which then run for its whole plan). The ungated figure puts
    every wave that gains on the network, whatever its hold-out error, and states the CVA errors
    that result (book total and worst netting set) against the plain FP64 run."""
    engine.set_shared_workspace(True)
    p32 = [engine.compile(s, abi.CCR_FP32) for s in specs]

    def wave_ms(run):
        run()
        engine.synchronize()
        t = []
        for _ in range(steps):
            run()
            t.append(engine.timing())
        return float(np.median([x.total_ms for x in t])), float(np.median([x.amc_ms for x in t]))

    reg = [wave_ms(p.execute_async) for p in p32]
    reg_out = [p.execute() for p in p32]
    engine.set_shared_workspace(False)
    surs, train_s = [], 0.0
    for s in specs:
        if not any(t.kind == abi.CCR_TRADE_BERMUDAN_SWAPTION for ns in s.sets for t in ns.trades):
            surs.append(None)
            continue
        mk = lambda seed: abi.Spec(s.model, s.sets, s.base_grid, 8000, seed, True, s.pfe_quantile, s.name)  # noqa: E731
        tr, ho = engine.compile(mk(101)), engine.compile(mk(202))
        t0 = time.perf_counter()
        surs.append(engine.surrogate_train(tr, ho, train_paths=4000, epochs=60, lr=5e-4, batch=4096, gate=1.0))
        train_s += time.perf_counter() - t0
        tr.close()
        ho.close()
    sur = [wave_ms(lambda: p.execute_surrogate_async(u)) if u else None for p, u in zip(p32, surs)]
    sur_out = [p.execute_surrogate(u) if u else None for p, u in zip(p32, surs)]
    errs = [[u.set_info(i)[1] for i in range(len(s.sets)) if u.set_info(i)[0] >= 0] if u else [] for s, u in zip(specs, surs)]

    def pick(gated: bool):
        use = [bool(u) and sm[0] < rm[0] and (not gated or max(e) <= gate)
               for u, sm, rm, e in zip(surs, sur, reg, errs)]
        ms = sum(sm[0] if on else rm[0] for on, sm, rm in zip(use, sur, reg))
        outs = [o for on, so, ro in zip(use, sur_out, reg_out) for o in (so if on else ro)]
        cva = np.array([float(o.get("cva_increments")[1:].sum()) for o in outs])
        ref = np.array([float(o.get("cva_increments")[1:].sum()) for o in outs64])
        rel = np.abs(cva - ref) / np.maximum(np.abs(ref), 1e-300)
        return {"fp32_surrogate_ms": ms, "speedup_vs_plain_fp64": ms64 / ms,
                "waves_on_surrogate": f"{sum(use)}/{len(use)}",
                "cva_rel_err": abs(cva.sum() - ref.sum()) / abs(ref.sum()),
                "worst_set_cva_rel_err": float(rel[ref > 1e-9 * np.abs(ref).max()].max())}

    all_errs = [e for w in errs for e in w]
    res = {"workload": "the benchmark's portfolio book (500 netting sets, 200,000 paths), waves as above",
           "plain_fp64_ms": ms64, "fp32_regression_ms": ms32,
           "gated": pick(True), "ungated": pick(False), "gate_ee": gate,
           "per_wave_ms": [{"regressions": round(rm[0], 2), "bermudan_stage_regressions": round(rm[1], 2),
                            "surrogate": round(sm[0], 2) if sm else None,
                            "bermudan_stage_surrogate": round(sm[1], 2) if sm else None,
                            "worst_holdout_ee_err": round(max(e), 4) if e else None}
                           for rm, sm, e in zip(reg, sur, errs)],
           "holdout_ee_err_median": float(np.median(all_errs)), "holdout_ee_err_max": float(np.max(all_errs)),
           "sets_within_gate": f"{sum(e <= gate for e in all_errs)}/{len(all_errs)}",
           "networks": sum(u.report.networks for u in surs if u), "train_s": train_s}
    res["fp32_surrogate_ms"] = res["gated"]["fp32_surrogate_ms"]
    res["speedup_surrogate_vs_plain"] = res["gated"]["speedup_vs_plain_fp64"]
    res["cva_rel_err"] = res["gated"]["cva_rel_err"]
    for p in p32:
        p.close()
    for u in surs:
        if u:
            u.close()
    return res


def surrogate_stage(paths: int, epochs: int, steps: int) -> dict:
    """The Bermudan stage (states pass plus regressions, or plus the network) of one FP32 execute of
    the portfolio book's first wave -- the 50 largest netting sets on the whole book's grids -- with
    the regression method and with the surrogate network (gate open), HIP-event timed."""
    engine = native.Engine()
    spec = W.c5(num_paths=paths)
    engine.set_book(True)
    rep, sim = native.book_grids(spec)
    engine.set_book_grids(rep, sim)
    wave = split_waves(spec, 50)[0]
    mk = lambda p, seed: abi.Spec(wave.model, wave.sets, wave.base_grid, p, seed, True, wave.pfe_quantile,  # noqa: E731
                                  wave.name)
    tr, ho = engine.compile(mk(8000, 101)), engine.compile(mk(8000, 202))
    print("surrogate stage: training plans compiled", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sur = engine.surrogate_train(tr, ho, train_paths=4000, epochs=epochs, lr=5e-4, batch=4096, gate=1.0)
    train_s = time.perf_counter() - t0
    print(f"surrogate stage: trained in {train_s:.1f} s", file=sys.stderr, flush=True)
    tr.close()
    ho.close()
    plan = engine.compile(mk(paths, 42), abi.CCR_FP32)
    print("surrogate stage: production plan compiled", file=sys.stderr, flush=True)

    def stage(run):
        run()
        engine.synchronize()
        t = []
        for _ in range(max(1, steps)):
            run()
            t.append(engine.timing())
        return float(np.median([x.amc_ms for x in t])), float(np.median([x.total_ms for x in t]))

    reg, reg_total = stage(plan.execute_async)
    net, net_total = stage(lambda: plan.execute_surrogate_async(sur))
    r = sur.report
    out = {"workload": "first wave of the portfolio book (50 largest netting sets), FP32 plan", "paths": paths,
           "bermudan_stage_regressions_ms": reg, "bermudan_stage_surrogate_ms": net, "stage_speedup": reg / net,
           "execute_regressions_ms": reg_total, "execute_surrogate_ms": net_total, "networks": r.networks,
           "outputs": r.candidate_sets, "holdout_ee_err_max": r.max_ee_error, "train_s": train_s, "epochs": epochs}
    plan.close()
    sur.close()
    engine.close()
    return out


def nested_im(workload: str, paths: int, inner: int, steps: int) -> dict:
    """GPU time of the nested conditional initial margin on one workload (the execute with it minus
    the plain execute, both HIP-event timed), and the FP64 oracle's time on a path sample."""
    if workload not in ("c1", "c2"):
        raise SystemExit("--nested-im runs on c1 or c2 (no Bermudans)")
    make, default_paths = WORKLOADS[workload]
    spec = make(paths or default_paths)
    q, mpor = 0.99, 10.0 / 250.0
    engine = native.Engine()

    def timed(plan):
        plan.execute_async()
        engine.synchronize()
        t = []
        for _ in range(max(1, steps)):
            plan.execute_async()
            t.append(engine.timing().total_ms)
        return float(np.median(t))

    plain = engine.compile(spec)
    plain_ms = timed(plain)
    plain.close()
    engine.set_nested_initial_margin(inner, q, mpor)
    plan = engine.compile(spec)
    nested_ms = timed(plan)
    info = plan.info()
    plan.execute()
    im = plan.nested_initial_margin(0)
    plan.close()
    engine.close()
    from tests import oracle_lib
    sample = make(64)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ref = oracle_lib.oracle_nested_im(sample, 0, inner, q, mpor, threads=threads)
    cpu_s = time.perf_counter() - t0
    reps = int(info.max_reporting_dates)
    evals = int(info.num_paths) * reps * inner * len(spec.sets)
    return {"workload": WORKLOAD_TEXT[workload], "paths": int(info.num_paths), "inner_paths": inner, "quantile": q,
            "execute_plain_ms": plain_ms, "execute_with_nested_im_ms": nested_ms,
            "nested_im_ms": nested_ms - plain_ms, "inner_revaluations": evals,
            "inner_revaluations_per_s": evals / ((nested_ms - plain_ms) * 1e-3),
            "set0_mean_im_first": float(im["mean"][0]),
            "cpu_oracle": {"sample": "netting set 0, 64 paths", "seconds": cpu_s, "threads": threads,
                           "inner_revaluations_per_s": 64 * len(ref["mean"]) * inner / cpu_s}}


def host_cpu() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return f"{line.split(':', 1)[1].strip()}, {os.cpu_count()} logical CPUs"
    except OSError:
        pass
    return f"{os.cpu_count()} logical CPUs"


def cpu_baseline(spec_fn, budget_s: float, threads: int, book: bool):
    """FP64 oracle on a bounded path sample of the same workload (all netting sets; the book
    restatement for book runs): all host threads (the reported value), then one thread on a
    smaller sample."""
    from tests import oracle_lib
    run = oracle_lib.run_oracle_book if book else oracle_lib.run_oracle

    def timed(nthreads: int, budget: float, probe_paths: int):
        probe = spec_fn(probe_paths)
        t0 = time.perf_counter()
        run(probe, threads=nthreads)
        per_path = (time.perf_counter() - t0) / probe_paths
        paths = int(max(probe_paths, min(2000, budget / max(per_path, 1e-9))))
        paths -= paths % 2
        spec = spec_fn(paths)
        t0 = time.perf_counter()
        run(spec, threads=nthreads)
        wall = time.perf_counter() - t0
        return spec, paths, wall, oracle_lib.oracle_timing()

    spec, paths, wall, tm = timed(threads, budget_s, 8)
    _, paths1, wall1, tm1 = timed(1, budget_s / 3, 2)
    what = "book restatement" if book else "ExposureEngine restated"
    return {
        "value": tm.revaluations / wall,
        "unit": "revals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle ({what}), same {len(spec.sets)} netting sets and grids, {paths} paths, "
                  f"{wall:.1f} s wall (draws {tm.draws_s:.2f} s, pricing {tm.pricing_s:.1f} s, "
                  f"collateral {tm.collateral_s:.2f} s, statistics {tm.statistics_s:.2f} s); host {host_cpu()}",
        "value_1thread": tm1.revaluations / wall1,
        "sample_1thread": f"same netting sets, {paths1} paths, {wall1:.1f} s wall on one thread",
    }


def compile_waves(engine, spec: abi.Spec, prec: int, wave_sets: int, book: bool):
    """One plan per wave of netting sets; book runs use the whole book's grids in every wave."""
    specs = split_waves(spec, wave_sets)
    engine.set_book(book)
    if book and len(specs) > 1:
        rep, sim = native.book_grids(spec)
        engine.set_book_grids(rep, sim)
    if len(specs) > 1:
        engine.set_shared_workspace(True)
    plans = [engine.compile(s, prec) for s in specs]
    return specs, plans


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--paths", type=int, default=0, help="0: the workload's own path count")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c5",
                    help="c5 (default: the largest single-GPU configuration); c1, c2 (3-ccy CSA sets), c3 "
                         "(Bermudan AMC), c4 (credit + Bermudans)")
    ap.add_argument("--book-group", type=int, default=0,
                    help="netting sets per workgroup of the book kernel (0: the library's default)")
    ap.add_argument("--block", type=int, default=0,
                    help="threads per workgroup of the path kernels (0: the library's occupancy choice)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="c5: keep the library's default book-kernel shape instead of autotuning it on the first wave")
    ap.add_argument("--no-book", action="store_true",
                    help="c5: independent netting sets (one engine run per set) instead of one book")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cva", action="store_true", help="skip the plain vs surrogate CVA wall time (c4)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 2^27 doubles on the GPU (1 GiB read, 1 GiB written): the reference launch "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--basis", choices=["default", "shared", "state", "sa"], default="default",
                    help="Bermudan regressions: shared (the state basis on shared per-date designs), state (the "
                         "same basis per record on the regression engine), sa (swap rate / annuity, total degree "
                         "3); default: sa for c3 (the configuration's basis), shared otherwise")
    ap.add_argument("--gram", choices=["fp64", "fp32"], default="fp32",
                    help="regression-engine Grams in FP64 or on FP32 matrix cores")
    ap.add_argument("--surrogate-stage", action="store_true",
                    help="only the Bermudan stage of the portfolio book's first wave (the 50 largest netting sets) "
                         "at --paths (default 200,000), FP32 plan: regressions against the surrogate network "
                         "(trained for --sur-epochs on 8,000-path plans)")
    ap.add_argument("--sur-epochs", type=int, default=60)
    ap.add_argument("--nested-im", type=int, default=0, metavar="INNER",
                    help="only the nested conditional initial margin (INNER inner paths, 99%%, 10-day margin period) "
                         "on --workload (c1 or c2) at --paths: the execute with and without it, and the FP64 oracle "
                         "on a path sample")
    ap.add_argument("--wave-sets", type=int, default=0,
                    help="netting sets per plan (0: all in one plan unless sets x paths exceeds 1e7)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path-sharded multi-GPU design is described, not built)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": src.numel() * 8, "bytes_written": dst.numel() * 8}))
        return

    if args.nested_im:
        print(json.dumps(nested_im(args.workload, args.paths, args.nested_im, args.steps)), flush=True)
        return
    if args.surrogate_stage:
        print(json.dumps(surrogate_stage(args.paths or 200000, args.sur_epochs, args.steps)), flush=True)
        return

    make, default_paths = WORKLOADS[args.workload]
    spec_fn = lambda p: make(p)  # noqa: E731
    spec = spec_fn(args.paths or default_paths)
    book = args.workload == "c5" and not args.no_book
    # Netting-set waves: portfolios whose per-execute buffers would not fit in HBM at once run as
    # one plan per wave, the waves sharing one context workspace (sized to the largest wave).
    wave_sets = args.wave_sets
    if wave_sets == 0 and len(spec.sets) * spec.num_paths > 10_000_000:
        wave_sets = max(1, 10_000_000 // spec.num_paths)
    prec = abi.CCR_FP32 if args.precision == "fp32" else abi.CCR_FP64
    engine = native.Engine(block_size=args.block)
    if args.book_group:
        engine.set_book_group(args.book_group)
    basis = args.basis if args.basis != "default" else ("sa" if args.workload == "c3" else "shared")
    gram = abi.CCR_FP32 if args.gram == "fp32" else abi.CCR_FP64
    engine.set_regression({"shared": -1, "state": abi.CCR_BASIS_STATE, "sa": abi.CCR_BASIS_SWAP_RATE_ANNUITY}[basis],
                          gram)
    tuned = None
    if book and not args.no_autotune and not args.book_group and not args.block:
        # The project's autotuner: (workgroup size x netting sets per workgroup) of the book kernel,
        # timed on the first wave (the largest sets) at the workload's own path count in this
        # precision: on a smaller sample the last, partly filled round of workgroups weighs more
        # and favours other shapes (65,536 paths: 2 sets per workgroup 1.4% ahead of 4; 200,000
        # paths: 4 sets 6% ahead of 2).
        engine.set_book(True)
        rep_grid, sim_grid = native.book_grids(spec)
        engine.set_book_grids(rep_grid, sim_grid)
        first = split_waves(spec, wave_sets)[0]
        t0 = time.perf_counter()
        blk, grp, tt = engine.autotune_book(first, prec, first.num_paths)
        tuned = {"block": blk, "sets_per_workgroup": grp, "seconds": time.perf_counter() - t0,
                 f"path_kernel_ms_{first.num_paths}_paths": {f"{b}x{g}": round(float(tt[i][k]), 3)
                                                for i, b in enumerate((64, 128, 192, 256))
                                                for k, g in enumerate((1, 2, 4, 8)) if tt[i][k] > 0}}
    specs, plans = compile_waves(engine, spec, prec, wave_sets, book)
    infos = [p.info() for p in plans]
    revals = sum(i.revaluations for i in infos)
    wall, kms, last = time_gpu(engine, plans, args.steps, args.warmup)
    ms_per_step = 1000.0 * wall / args.steps
    value = revals / (wall / args.steps)

    elem = 4 if prec == abi.CCR_FP32 else 8
    outs = [o for p in plans for o in p.execute()]
    path_bytes = path_kernel_bytes(plans, outs, infos, elem)
    achieved_gbs = path_bytes / (kms["path"] * 1e-3) / 1e9
    variant = f"{args.workload}{'' if book or args.workload != 'c5' else '_sets'}"

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "revals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.CCR_FP64 else "f32",
        "data": "synthetic (seeded SplitMix64 portfolio, replayed std::mt19937_64 normal stream)",
        "config": describe(spec, infos, len(specs)),
        "kernel_ms": kms,
        "host_ms": {"draws": last.draws_ms, "upload": last.upload_ms},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes": path_bytes,
            "kernel": ("ccr_group_kernel" if book else "ccr_path_kernel") + " (all waves of one step)",
        },
    }
    if book:
        result["config"]["book"] = True
    if tuned:
        result["autotune"] = tuned
    if any(i.num_bermudans for i in infos):
        result["config"]["regression"] = {
            "basis": {"shared": "state (1, z, z^2, z^3) on shared per-date designs",
                      "state": "state (1, z, z^2, z^3) per record",
                      "sa": "swap rate / annuity, total degree 3 (k = 10) per record"}[basis]}
        if basis != "shared":
            result["config"]["regression"]["gram"] = "fp32 matrix cores" if gram == abi.CCR_FP32 else "fp64"
    traffic, source = measured_traffic(variant, args.precision)
    if traffic is not None and not args.paths:
        result["roofline"]["traffic"] = traffic
        result["roofline"]["traffic_source"] = source
        result["roofline"]["reread_factor"] = traffic / path_bytes
    # The ceiling that binds this kernel: VALU issue. Every wave64 VALU instruction occupies its
    # SIMD for 4 cycles in FP64 (16 lanes per cycle) and 2 in FP32; 1,024 SIMDs at 2.4 GHz.
    valu, vsource = profile_figure(variant, args.precision, "valu_instructions")
    if valu is not None and not args.paths:
        cycles = 4.0 if prec == abi.CCR_FP64 else 2.0
        issue = valu * cycles / (VALU_SIMDS * VALU_CLOCK_HZ * kms["path"] * 1e-3)
        result["compute"] = {"bound": "valu", "valu_instructions": valu, "issue_frac": issue,
                             "source": vsource}
    for p in plans:
        p.close()
    if not args.no_fp32 and prec == abi.CCR_FP64:
        _, p32 = compile_waves(engine, spec, abi.CCR_FP32, wave_sets, book)
        w32, k32, _ = time_gpu(engine, p32, args.steps, args.warmup)
        o32 = [o for p in p32 for o in p.execute()]
        for p in p32:
            p.close()
        ee_err = max(float(np.abs(a.get("expected_exposure") - b.get("expected_exposure")).max()
                           / np.abs(b.get("expected_exposure")).max()) for a, b in zip(o32, outs))
        cva64 = sum(float(o.get("cva_increments")[1:].sum()) for o in outs)
        cva32 = sum(float(o.get("cva_increments")[1:].sum()) for o in o32)
        result["fp32"] = {"value": revals / (w32 / args.steps), "ms_per_step": 1000.0 * w32 / args.steps,
                          "gpu_ms_per_step": k32["total"], "path_kernel_ms": k32["path"], "ee_rel_err_vs_fp64": ee_err,
                          "cva_fp64": cva64, "cva_fp32": cva32,
                          "cva_rel_err_vs_fp64": abs(cva32 - cva64) / abs(cva64) if cva64 else None}
    if not args.no_cva and args.workload == "c5" and book and prec == abi.CCR_FP64 and not args.no_fp32:
        # GPU time per CVA of the whole book (HIP events around each wave's execute, summed).
        result["cva_wall_time"] = cva_wall_time_c5(engine, specs, args.steps, outs, kms["total"],
                                                   result["fp32"]["gpu_ms_per_step"])
        result["cva_wall_time"]["c4"] = cva_wall_time_c4(engine, args.steps)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(spec_fn, args.cpu_seconds, threads, book)
    engine.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
