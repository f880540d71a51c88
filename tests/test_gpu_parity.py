"""HIP exposure path vs the FP64 oracle on the same seeded inputs (same replayed normal stream).

FP64 bar: every profile field within 1e-10 scale-relative (see compare.py), reporting dates
bit-identical. FP32 runs state their error against the FP64 oracle.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import compare, oracle_lib
from zlatko_amd import abi
from zlatko_amd import workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10

SMALL = {
    "swaps": dict(),
    "csa": dict(csa=True),
    "fx": dict(fx=True),
    "credits": dict(credits=True),
    "two_leg": dict(two_leg=True),
    "all": dict(csa=True, fx=True, credits=True, two_leg=True),
    "options": dict(fx=True, csa=True, credits=True, options=6),
    "csa_bond": dict(csa=True, credits=True, security="bond"),
    "csa_asset": dict(csa=True, fx=True, credits=True, security="asset"),
}


def _check(spec, outs, ref, tol=FP64_TOL, pointwise=True):
    for s in range(len(spec.sets)):
        errs = compare.field_errors(outs[s], ref[s])
        print(spec.name, s, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert errs["times"] == 0.0
        assert compare.worst(errs) <= tol, errs
        if pointwise:
            pe = compare.pointwise_errors(outs[s], ref[s])
            assert compare.worst(pe) <= tol, pe


@pytest.mark.parametrize("case", sorted(SMALL))
def test_small_cases_fp64(engine, case):
    spec = W.small_reference_case(num_paths=256, **SMALL[case])
    _check(spec, engine.run(spec), oracle_lib.run_oracle(spec, threads=4))


GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


@pytest.mark.parametrize("case", sorted(GOLDEN["cases"]))
def test_reference_golden_vectors_fp64(engine, case):
    """The profiles a build of the reference engine produced on the same inputs (swaps, CSA, FX
    forwards, CIR++ credits, two-leg swaps, Bermudans physical and cash, FX options)."""
    g = GOLDEN["cases"][case]
    spec = W.small_reference_case(**g["params"])
    out = engine.run(spec)[0]
    scale = max(np.abs(np.array(g["profile"]["expected_exposure"])).max(), 1.0)
    for f, v in g["profile"].items():
        err = np.abs(out.get(f) - np.array(v)).max() / (1.0 if f == "times" else scale)
        print(case, f, f"{err:.2e}", flush=True)
        assert err <= (0.0 if f == "times" else FP64_TOL), f
    if "pathwise_cva" in g and spec.sets[0].counterparty_credit >= 0:
        cva = out.get("cva_increments")[1:].sum()
        assert abs(cva - g["pathwise_cva"]) <= FP64_TOL * max(1.0, abs(g["pathwise_cva"]))


def test_c1_full_size_fp64(engine):
    """Config 1 at its full size: 10 two-leg IRS, HW1F, 5,000 paths x 120 monthly dates, PFE 97.5%."""
    spec = W.c1()
    _check(spec, engine.run(spec), oracle_lib.run_oracle(spec, threads=8))


def test_c2_reduced_fp64(engine):
    """3-currency HW + 2 FX, CSA netting sets (config 2 shape, fewer paths/sets)."""
    spec = W.c2(num_paths=1000, num_sets=4, num_trades=80, weeks=104, horizon=2.0)
    _check(spec, engine.run(spec), oracle_lib.run_oracle(spec, threads=8))


def test_c2_full_portfolio_fp64(engine):
    """Config 2 in full -- 20 CSA netting sets, 1,000 trades (700 two-leg IRS in 3 currencies, 300
    FX forwards), 260 weekly reporting dates (up to 619 simulation dates with the margin-call and
    event dates), the plan shape the path kernel is tuned for -- at 384 paths: every set and date
    against the oracle, scale-relative and pointwise."""
    spec = W.c2(num_paths=384)
    plan = engine.compile(spec)
    info = plan.info()
    outs = plan.execute()
    plan.close()
    assert info.num_sets == 20 and info.max_sim_dates > 600
    ref = oracle_lib.run_oracle(spec, threads=16)
    _check(spec, outs, ref)


def test_c4_plain_credit_fp64(engine):
    """CIR++ counterparty and own intensities with wrong-way correlation: CVA/DVA increments."""
    spec = W.c4(num_paths=2000, dates=60, with_bermudans=False)
    outs = engine.run(spec)
    ref = oracle_lib.run_oracle(spec, threads=8)
    _check(spec, outs, ref)
    assert outs[0].get("cva_increments")[1:].sum() > 0


def test_c5_uneven_portfolio_fp64(engine):
    """Config 5 shape: 500 netting sets with Zipf(1.2) trade counts (1 .. 5,000; ~20,700 trades:
    IRS in 3 currencies, FX forwards, ~1,000 Bermudans), half under CSA, shared CIR++ credit, all
    in one plan and one launch per stage; every set against the oracle."""
    spec = W.c5(num_paths=64)
    outs = engine.run(spec)
    ref = oracle_lib.run_oracle(spec, threads=16)
    worst = max(compare.worst(compare.field_errors(a, b)) for a, b in zip(outs, ref))
    print("c5", len(spec.sets), "sets, worst field error", f"{worst:.2e}", flush=True)
    assert worst <= FP64_TOL


def test_autotuner_picks_a_size_and_keeps_results(engine):
    from zlatko_amd import native
    spec = W.c2(num_paths=4000, num_sets=4, num_trades=80, weeks=52, horizon=1.0)
    e = native.Engine()
    before = e.run(spec)[0].as_dict()
    block, times = e.autotune(spec, sample_paths=4000)
    print("autotune", block, times, flush=True)
    assert block in (64, 128, 192, 256) and np.all(times > 0)
    after = e.run(spec)[0].as_dict()
    e.close()
    for k in abi.PROFILE_FIELDS:
        assert np.array_equal(before[k], after[k]), k


def test_book_autotuner_picks_a_shape_and_keeps_results():
    """The book kernel's autotuner over (workgroup size x sets per workgroup): a measured choice,
    and the book's results are bit-identical before and after (every shape sums in one order)."""
    from zlatko_amd import native
    spec = W.c5(num_paths=2048, num_sets=12, largest=60, dates=24)
    e = native.Engine()
    try:
        e.set_book(True)

        def run():
            plan = e.compile(spec)
            outs = [o.as_dict() for o in plan.execute()]
            plan.close()
            return outs

        before = run()
        block, group, times = e.autotune_book(spec, sample_paths=2048)
        print("book autotune", block, group, times.tolist(), flush=True)
        assert block in (64, 128, 192, 256) and group in (1, 2, 4, 8)
        assert times[(64, 128, 192, 256).index(block)][(1, 2, 4, 8).index(group)] == times[times > 0].min()
        assert (times > 0).sum() >= 8
        after = run()
        for a, b in zip(before, after):
            for k in abi.PROFILE_FIELDS:
                assert np.array_equal(a[k], b[k]), k
    finally:
        e.close()


def test_fp32_error_vs_fp64_oracle(engine):
    spec = W.c2(num_paths=1000, num_sets=2, num_trades=40, weeks=52, horizon=1.0)
    ref = oracle_lib.run_oracle(spec, threads=8)
    outs = engine.run(spec, abi.CCR_FP32)
    for s in range(len(spec.sets)):
        errs = compare.field_errors(outs[s], ref[s])
        print("fp32", s, {k: f"{v:.2e}" for k, v in errs.items()})
        assert errs["times"] == 0.0
        assert errs["expected_exposure"] < 1e-4 and errs["discounted_expected_exposure"] < 1e-4


def test_cpp_wrapper_run_matches_oracle(tmp_path):
    """ccr::hip::ExposureEngine::run (include/ccr_hip.hpp) through tests/cpp/wrapper_smoke.cpp:
    EE and EPE against the FP64 oracle on the same inputs."""
    import subprocess
    from tests.test_abi import build_wrapper_smoke
    out = subprocess.run([str(build_wrapper_smoke(tmp_path))], capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()
    assert not out.startswith("error"), out
    t = np.array([float(l.split()[0]) for l in lines[:-1]])
    ee = np.array([float(l.split()[1]) for l in lines[:-1]])
    epe = float(lines[-1].split()[1])
    model = abi.Model(abi.HullWhite(abi.Curve([1.0, 3.0, 10.0], [0.02, 0.025, 0.03]), 0.05, 0.01),
                      assets=[abi.Asset(1.1, 0.12, abi.Curve.flat(0.01))],
                      correlation=np.array([[1.0, 0.3], [0.3, 1.0]]))
    trades = [abi.Trade(abi.CCR_TRADE_SWAP, 1e6, 0.025, True, schedule=[0.0, 0.5, 1.0, 1.5, 2.0]),
              abi.Trade(abi.CCR_TRADE_SWAP, 5e5, 0.03, False, schedule=[0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0]),
              abi.Trade(abi.CCR_TRADE_ASSET_FORWARD, 4e5, 1.1, asset=0, maturity=2.5)]
    spec = abi.Spec(model, [abi.NettingSet(trades)], np.array([0.25 * i for i in range(13)]), 2000, 42, True, 0.95,
                    "wrapper")
    ref = oracle_lib.run_oracle(spec)[0]
    assert np.array_equal(t, ref.get("times"))
    scale = np.abs(ref.get("expected_exposure")).max()
    assert np.abs(ee - ref.get("expected_exposure")).max() <= FP64_TOL * scale
    ref_epe = oracle_lib.oracle().ccr_oracle_epe(abi.dptr(t), abi.dptr(ref.get("expected_exposure")), len(t), 3.0)
    assert abs(epe - ref_epe) <= FP64_TOL * scale


def test_fp32_options_error_vs_fp64_oracle(engine):
    """FX options (Black in FP32: logf, erfcf) next to swaps and forwards."""
    spec = W.small_reference_case(num_paths=2000, fx=True, csa=True, credits=True, options=6)
    ref = oracle_lib.run_oracle(spec, threads=8)
    out = engine.run(spec, abi.CCR_FP32)
    errs = compare.field_errors(out[0], ref[0])
    print("fp32 options", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert errs["expected_exposure"] < 1e-4 and errs["discounted_expected_exposure"] < 1e-4


@pytest.mark.parametrize("paths,precision,level", [
    (1, abi.CCR_FP64, None), (2, abi.CCR_FP64, None), (3, abi.CCR_FP64, None), (777, abi.CCR_FP64, None),
    (2049, abi.CCR_FP64, None), (4096, abi.CCR_FP64, None), (4097, abi.CCR_FP64, None), (5001, abi.CCR_FP64, None),
    (20000, abi.CCR_FP64, 0.999), (20000, abi.CCR_FP64, 0.001), (20000, abi.CCR_FP64, 0.3),
    (70000, abi.CCR_FP64, 0.5), (120000, abi.CCR_FP64, None), (120000, abi.CCR_FP64, 0.5),
    (777, abi.CCR_FP32, None), (5001, abi.CCR_FP32, None), (70000, abi.CCR_FP32, 0.5), (120000, abi.CCR_FP32, None)])
def test_pfe_is_exact_order_statistic(engine, paths, precision, level):
    """The GPU PFE equals ccr::quantile of the GPU's own exposure rows, bit for bit. Rows include
    all-equal values (t = 0: ties at the splitters), all-zero rows (after the last cash flow),
    quantiles whose sample splitter falls off either end of the sample, P = 70,000 at the median,
    where more keys lie between the splitters than LDS holds (the radix-select path), and P =
    120,000 (the 8,192-key buffer)."""
    spec = W.small_reference_case(num_paths=paths, csa=True, fx=True)
    if level is not None:
        spec.pfe_quantile = level
    out = engine.run(spec, precision, rows=True)[0]
    rows = out.exposure_rows(spec.num_paths)
    pfe = out.get("potential_future_exposure")
    for q_level in (spec.pfe_quantile,):
        for r in range(rows.shape[0]):
            v = np.sort(rows[r])
            if precision == abi.CCR_FP32:
                v = v.astype(np.float32).astype(np.float64)
            h = q_level * (len(v) - 1)
            lo = int(np.floor(h))
            hi = min(lo + 1, len(v) - 1)
            expect = v[lo] + (h - lo) * (v[hi] - v[lo])
            assert pfe[r] == expect, (r, pfe[r], expect)


def test_bit_reproducible_across_runs_and_block_sizes():
    from zlatko_amd import native
    spec = W.small_reference_case(num_paths=1000, csa=True, credits=True)
    results = []
    for block in (64, 128, 192, 256):
        e = native.Engine(block_size=block)
        results.append(e.run(spec)[0].as_dict())
        results.append(e.run(spec)[0].as_dict())
        e.close()
    for r in results[1:]:
        for k in abi.PROFILE_FIELDS:
            assert np.array_equal(r[k], results[0][k]), k


def test_plan_execute_matches_run(engine):
    spec = W.small_reference_case(num_paths=300, fx=True)
    plan = engine.compile(spec)
    a = plan.execute()[0].as_dict()
    b = plan.execute()[0].as_dict()
    c = engine.run(spec)[0].as_dict()
    info = plan.info()
    plan.close()
    assert info.num_paths == 300 and info.revaluations > 0
    for k in abi.PROFILE_FIELDS:
        assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], c[k]), k


def test_forced_failures_leave_engine_usable(engine):
    from zlatko_amd import native
    spec = W.small_reference_case(num_paths=64)
    for fault, code in ((abi.CCR_FAULT_ALLOC, abi.CCR_ERR_OUT_OF_MEMORY), (abi.CCR_FAULT_LAUNCH, abi.CCR_ERR_DEVICE),
                        (abi.CCR_FAULT_PLAN, abi.CCR_ERR_INVALID_ARGUMENT)):
        engine.set_fault(fault)
        with pytest.raises(native.CcrError) as e:
            engine.run(spec)
        assert e.value.code == code
    _check(spec, engine.run(spec), oracle_lib.run_oracle(spec))


def test_plans_sharing_a_context_stream_match_fresh_contexts():
    """Plans of one context replay prefixes of one device stream per seed (generated once; FP32
    plans narrow it on the device). Interleaved compiles of different lengths, seeds and
    precisions give the same profiles as plans compiled on fresh contexts."""
    from zlatko_amd import native
    short = W.small_reference_case(num_paths=300, csa=True, fx=True)
    long_ = W.small_reference_case(num_paths=900, csa=True, fx=True, bermudans=2)
    other = W.small_reference_case(num_paths=300, csa=True, fx=True)
    other.config.seed = 7
    order = [(short, abi.CCR_FP64), (long_, abi.CCR_FP32), (other, abi.CCR_FP64), (short, abi.CCR_FP32),
             (long_, abi.CCR_FP64), (short, abi.CCR_FP64)]
    e = native.Engine()
    plans = [e.compile(s, p) for s, p in order]
    shared = [pl.execute()[0].as_dict() for pl in plans]
    for pl in plans:
        pl.close()
    e.close()
    for (s, p), got in zip(order, shared):
        f = native.Engine()
        want = f.run(s, p)[0].as_dict()
        f.close()
        for k in abi.PROFILE_FIELDS:
            assert np.array_equal(got[k], want[k]), (s.num_paths, p, k)


def test_shared_workspace_waves_match_owned_buffers():
    """Waves of netting sets as plans sharing one context workspace (grown when a larger wave
    executes) give the profiles of plans that own their buffers; a surrogate refuses them."""
    from zlatko_amd import native
    a = W.small_reference_case(num_paths=500, csa=True, fx=True, credits=True, bermudans=2)
    b = W.c5(num_paths=400)
    e = native.Engine()
    e.set_shared_workspace(True)
    pa, pb = e.compile(a), e.compile(b)
    got = [pa.execute(), pb.execute(), pa.execute()]
    with pytest.raises(native.CcrError):
        e.surrogate_train(pa, pb, train_paths=64, epochs=1)
    pa.close()
    pb.close()
    e.close()
    f = native.Engine()
    want_a, want_b = f.run(a), f.run(b)
    f.close()
    for outs, want in ((got[0], want_a), (got[1], want_b), (got[2], want_a)):
        for o, w in zip(outs, want):
            for k in abi.PROFILE_FIELDS:
                assert np.array_equal(o.get(k), w.get(k)), k


IM_CASES = {
    "csa_fx_credits": dict(csa=True, fx=True, credits=True),
    "two_leg": dict(two_leg=True),
    "bermudans": dict(csa=True, fx=True, bermudans=3),
    "options": dict(fx=True, credits=True, options=6),
}


def _im_check(plan, spec, quantile, mpor, tol):
    worst = 0.0
    for s in range(len(spec.sets)):
        got = plan.initial_margin(s)
        want = oracle_lib.oracle_initial_margin(spec, s, quantile, mpor)
        assert len(got) == len(want) and got[0] == 0.0
        err = np.abs(got - want).max() / max(np.abs(want).max(), 1.0)
        worst = max(worst, err)
        assert err <= tol, (s, err)
    return worst


@pytest.mark.parametrize("case", sorted(IM_CASES))
def test_initial_margin_matches_oracle_fp64(case):
    """initialMarginProfile (src/xva/xva.cpp:59-76) from the path kernel's |V(t_k) - V(t_k-1)|
    rows and per-wave deflator sums, against the oracle on the same paths (700 paths: several
    wavefronts per set); the exposure profile of the same execute is unchanged."""
    from zlatko_amd import native
    spec = W.small_reference_case(num_paths=700, **IM_CASES[case])
    e = native.Engine()
    e.set_initial_margin(0.99, 10.0 / 250.0)
    plan = e.compile(spec)
    _check(spec, plan.execute(), oracle_lib.run_oracle(spec, threads=4))
    print(case, f"{_im_check(plan, spec, 0.99, 10.0 / 250.0, FP64_TOL):.2e}", flush=True)
    plan.close()
    e.close()


def test_initial_margin_many_sets_fp64_and_fp32():
    """Several netting sets of uneven size in one plan (FP64 within 1e-10; the FP32 plan's error
    against the FP64 oracle is printed and bounded by 1e-4 of the largest margin)."""
    from zlatko_amd import native
    spec = W.c2(num_paths=1000, num_sets=4, num_trades=80, weeks=104, horizon=2.0)
    e = native.Engine()
    e.set_initial_margin(0.975, 0.05)
    p64, p32 = e.compile(spec, abi.CCR_FP64), e.compile(spec, abi.CCR_FP32)
    p64.execute()
    p32.execute()
    _im_check(p64, spec, 0.975, 0.05, FP64_TOL)
    print("fp32 initial margin error", f"{_im_check(p32, spec, 0.975, 0.05, 1e-4):.2e}", flush=True)
    p64.close()
    p32.close()
    e.close()


def test_initial_margin_errors():
    """A plan compiled without it, or not executed yet, refuses; quantile 0 switches it off."""
    from zlatko_amd import native
    spec = W.small_reference_case(num_paths=128, csa=True)
    e = native.Engine()
    plain = e.compile(spec)
    plain.execute()
    with pytest.raises(native.CcrError):
        plain.initial_margin(0)
    e.set_initial_margin(0.99)
    fresh = e.compile(spec)
    with pytest.raises(native.CcrError):
        fresh.initial_margin(0)
    with pytest.raises(native.CcrError):
        e.set_initial_margin(0.99, -1.0)
    fresh.execute()
    with pytest.raises(native.CcrError):
        fresh.initial_margin(1)
    assert fresh.initial_margin(0)[0] == 0.0
    e.set_initial_margin(0.0)
    off = e.compile(spec)
    off.execute()
    with pytest.raises(native.CcrError):
        off.initial_margin(0)
    for p in (plain, fresh, off):
        p.close()
    e.close()


TRADE_CASES = {
    "swaps": dict(),
    "csa_fx_credits": dict(csa=True, fx=True, credits=True),
    "options": dict(fx=True, credits=True, options=6),
    "two_leg": dict(two_leg=True, csa=True),
}


def _trade_check(plan, spec, tol):
    worst = 0.0
    for s in range(len(spec.sets)):
        _, want = oracle_lib.oracle_trade_statistics(spec, s)
        n = spec.sets_c[s].num_trades
        scale = max(np.abs(want["standalone"]).max(), 1.0)
        for i in range(n):
            got = plan.trade_statistics(s, i)
            for m in abi.TRADE_STATS:
                assert len(got[m]) == want[m].shape[1]
                err = np.abs(got[m] - want[m][i]).max() / scale
                worst = max(worst, err)
                assert err <= tol, (s, i, m, err)
    return worst


@pytest.mark.parametrize("case", sorted(TRADE_CASES))
def test_trade_statistics_match_oracle_fp64(case):
    """Per-trade mode (each trade valued on its own, netted in trade order): Euler contributions
    with and without the deflator, incremental EE* and standalone EE / EE* per trade against the
    oracle (allocation.cpp's estimators); the plan's exposure profile stays within 1e-10."""
    from zlatko_amd import native
    spec = W.small_reference_case(num_paths=700, **TRADE_CASES[case])
    e = native.Engine()
    e.set_trade_statistics(True)
    plan = e.compile(spec)
    _check(spec, plan.execute(), oracle_lib.run_oracle(spec, threads=4))
    print(case, f"{_trade_check(plan, spec, FP64_TOL):.2e}", flush=True)
    plan.close()
    e.close()


def test_trade_statistics_many_sets_fp64_and_fp32():
    """Several CSA netting sets with foreign swaps and FX forwards (FP32 error against the FP64
    oracle printed, bounded by 1e-4 of the largest standalone EE)."""
    from zlatko_amd import native
    spec = W.c2(num_paths=800, num_sets=3, num_trades=30, weeks=52, horizon=1.0)
    assert max(spec.sets_c[s].num_trades for s in range(3)) <= 16
    e = native.Engine()
    e.set_trade_statistics(True)
    p64, p32 = e.compile(spec, abi.CCR_FP64), e.compile(spec, abi.CCR_FP32)
    p64.execute()
    p32.execute()
    _trade_check(p64, spec, FP64_TOL)
    print("fp32 per-trade error", f"{_trade_check(p32, spec, 1e-4):.2e}", flush=True)
    p64.close()
    p32.close()
    e.close()


def test_trade_statistics_errors():
    """More than 16 trades or a Bermudan swaption refuse to compile (CCR_ERR_UNSUPPORTED); a plan
    compiled without per-trade mode, or not executed, refuses the query; indices are checked."""
    from zlatko_amd import native
    e = native.Engine()
    e.set_trade_statistics(True)
    for spec in (W.small_reference_case(num_paths=64, fx=True, options=11),  # 17 trades
                 W.small_reference_case(num_paths=64, bermudans=1)):
        with pytest.raises(native.CcrError) as err:
            e.compile(spec)
        assert err.value.code == abi.CCR_ERR_UNSUPPORTED
    spec = W.small_reference_case(num_paths=128)
    fresh = e.compile(spec)
    with pytest.raises(native.CcrError):
        fresh.trade_statistics(0, 0)
    fresh.execute()
    for s, i in ((0, 4), (1, 0), (0, -1)):
        with pytest.raises(native.CcrError) as err:
            fresh.trade_statistics(s, i)
        assert err.value.code == abi.CCR_ERR_OUT_OF_RANGE
    e.set_trade_statistics(False)
    off = e.compile(spec)
    off.execute()
    with pytest.raises(native.CcrError):
        off.trade_statistics(0, 0)
    for p in (fresh, off):
        p.close()
    e.close()


def test_xva_allocation_of_per_trade_contributions_adds_up():
    """xva_allocation_adds_up (tests/test_xva.cpp:90-106): CVA, FCA and KVA of each trade's
    discounted Euler contribution from the GPU add up to those of the netting set's EE* (no CSA)
    within 1e-10 of the total."""
    import ctypes as C
    from zlatko_amd import native
    spec = W.small_reference_case(num_paths=2000, fx=True, credits=True, options=4)
    e = native.Engine()
    e.set_trade_statistics(True)
    plan = e.compile(spec)
    prof = plan.execute()[0]
    n = spec.sets_c[0].num_trades
    contrib = np.ascontiguousarray([plan.trade_statistics(0, i)["marginal_discounted"] for i in range(n)])
    t = prof.get("times")
    piv, ch, oh = np.array([1.0, 3.0, 5.0]), np.array([0.01, 0.015, 0.02]), np.array([0.006, 0.006, 0.008])
    cpty, own = abi.CreditCurveC(abi.dptr(piv), abi.dptr(ch), 3), abi.CreditCurveC(abi.dptr(piv), abi.dptr(oh), 3)
    p = abi.XvaParametersC()
    native.lib().ccr_xva_defaults(C.byref(p))
    parts = np.zeros((n, 3))
    assert native.lib().ccr_allocate_exposure_xva(abi.dptr(t), abi.dptr(contrib), n, len(t), C.byref(cpty),
                                                  C.c_double(0.4), C.byref(own), C.byref(p), abi.dptr(parts)) == 0
    whole = np.zeros(3)
    assert native.lib().ccr_exposure_xva(abi.dptr(t), abi.dptr(prof.get("discounted_expected_exposure")), len(t),
                                         C.byref(cpty), C.c_double(0.4), C.byref(own), C.byref(p),
                                         abi.dptr(whole)) == 0
    assert (whole > 0).all()
    assert np.all(np.abs(parts.sum(axis=0) - whole) <= 1e-10 * whole), (parts.sum(axis=0), whole)
    plan.close()
    e.close()


@pytest.mark.parametrize("secured", [False, True])
def test_book_matches_oracle_and_totals_add_up_fp64(secured):
    """Portfolio book: every netting set on the union grids with pathwise identical scenarios
    (E[D] bit-identical across sets on the GPU), profiles within 1e-10 of the oracle's book, and
    ccr_book_totals per counterparty equal to the sums of the sets' additive fields. secured: two
    of the sets post a zero bond and the FX asset as collateral."""
    from zlatko_amd import native
    spec = W.book_case(num_paths=3000, secured=secured)
    e = native.Engine()
    e.set_book(True)
    plan = e.compile(spec)
    outs = plan.execute()
    _check(spec, outs, oracle_lib.run_oracle_book(spec, threads=4))
    for o in outs[1:]:
        assert np.array_equal(o.get("mean_deflator"), outs[0].get("mean_deflator"))
    arr = (abi.ExposureOutC * len(outs))(*[o.c for o in outs])
    for cp, members in ((0, [0, 1]), (1, [2]), (-1, [0, 1, 2])):
        tot, tarr = abi.make_outs(1, outs[0].c.num_dates, 0, False)
        assert native.lib().ccr_book_totals(arr, spec.sets_c, len(outs), cp, tarr) == 0
        abi.sync_outs(tot, tarr)
        for f in ("expected_exposure", "discounted_expected_exposure", "cva_increments", "dva_increments"):
            want = sum(outs[i].get(f) for i in members)
            assert np.abs(tot[0].get(f) - want).max() <= 1e-12 * max(np.abs(want).max(), 1.0), (cp, f)
    tot, tarr = abi.make_outs(1, outs[0].c.num_dates, 0, False)
    assert native.lib().ccr_book_totals(arr, spec.sets_c, len(outs), 5, tarr) == abi.CCR_ERR_OUT_OF_RANGE
    plan.close()
    e.set_book(False)
    single = e.compile(spec)
    assert single.info().max_reporting_dates < outs[0].c.num_dates  # own grids again
    single.close()
    e.close()


def _book_spec(num_paths):
    """A config-5-shaped book small enough for the oracle: 24 uneven netting sets (Zipf sizes
    up to 60 trades), IRS in 3 currencies, FX forwards and Bermudans (10% of trades), every other
    set under CSA, one shared CIR++ counterparty credit."""
    return W.c5(num_paths=num_paths, num_sets=24, largest=60, bermudan_share=0.1)


def test_book_with_bermudans_matches_oracle_fp64():
    """Book plans share one states pass and one regression design per date among every set with
    Bermudans, and tile the continuation GEMM across sets valued on the same dates: profiles of
    every set within 1e-10 of the oracle's book (scale-relative and pointwise)."""
    from zlatko_amd import native
    spec = _book_spec(512)
    e = native.Engine()
    e.set_book(True)
    plan = e.compile(spec)
    info = plan.info()
    outs = plan.execute()
    plan.close()
    e.close()
    holders = sum(1 for ns in spec.sets if any(t.kind == abi.CCR_TRADE_BERMUDAN_SWAPTION for t in ns.trades))
    print("book with Bermudans:", holders, "sets hold", info.num_bermudans, "swaptions", flush=True)
    assert holders >= 4
    _check(spec, outs, oracle_lib.run_oracle_book(spec, threads=16))


@pytest.mark.parametrize("precision", [abi.CCR_FP64, abi.CCR_FP32])
def test_book_kernel_group_sizes_give_identical_results(precision):
    """The book kernel values 1, 2, 4 or 8 netting sets per workgroup from one market step per
    date: every choice gives the same profiles bit for bit (FP64 and FP32), with 24 sets so the
    last group is partial for 8 ... and exact for the others."""
    from zlatko_amd import native
    spec = _book_spec(777)
    got = {}
    for group in (1, 2, 4, 8):
        e = native.Engine()
        e.set_book(True)
        e.set_book_group(group)
        p = e.compile(spec, precision)
        got[group] = p.execute(rows=True)
        p.close()
        e.close()
    for group in (2, 4, 8):
        for a, b in zip(got[group], got[1]):
            for f in abi.PROFILE_FIELDS:
                assert np.array_equal(a.get(f), b.get(f)), (group, f)
            assert np.array_equal(a.exposure_rows(spec.num_paths), b.exposure_rows(spec.num_paths)), group


@pytest.mark.parametrize("precision", [abi.CCR_FP64, abi.CCR_FP32])
def test_book_kernel_group_sizes_agree_on_security_collateral(precision):
    """Bond and FX-asset collateral in the book kernel: 1, 2, 4 or 8 sets per workgroup give the
    same profiles and exposure rows bit for bit."""
    from zlatko_amd import native
    spec = W.book_case(num_paths=513, secured=True)
    got = {}
    for group in (1, 2, 4, 8):
        e = native.Engine()
        e.set_book(True)
        e.set_book_group(group)
        p = e.compile(spec, precision)
        got[group] = p.execute(rows=True)
        p.close()
        e.close()
    for group in (2, 4, 8):
        for a, b in zip(got[group], got[1]):
            for f in abi.PROFILE_FIELDS:
                assert np.array_equal(a.get(f), b.get(f)), (group, f)
            assert np.array_equal(a.exposure_rows(spec.num_paths), b.exposure_rows(spec.num_paths)), group


def test_book_in_waves_matches_one_plan_fp64():
    """A book compiled as waves of netting sets on the whole book's grids (ccr_hip_book_grids,
    ccr_hip_set_book_grids; one shared workspace) reproduces the single book plan: same dates,
    identical deflators and survivals, profiles within 1e-13 (the continuation GEMM's tiles
    group the swaptions differently)."""
    from zlatko_amd import native
    spec = _book_spec(1024)
    e = native.Engine()
    e.set_book(True)
    whole = e.compile(spec)
    want = whole.execute()
    whole.close()
    rep, sim = native.book_grids(spec)
    e.set_book_grids(rep, sim)
    e.set_shared_workspace(True)
    got = []
    for w0 in range(0, len(spec.sets), 8):
        sub = abi.Spec(spec.model, spec.sets[w0:w0 + 8], spec.base_grid, spec.num_paths, int(spec.config.seed),
                       bool(spec.config.antithetic), spec.pfe_quantile, spec.name)
        p = e.compile(sub)
        got += p.execute()
        p.close()
    e.close()
    for a, b in zip(got, want):
        assert np.array_equal(a.get("times"), b.get("times"))
        assert np.array_equal(a.get("mean_deflator"), b.get("mean_deflator"))
        assert np.array_equal(a.get("mean_survival"), b.get("mean_survival"))
        errs = compare.field_errors(a, b)
        assert compare.worst(errs) <= 1e-13, errs


def test_book_initial_margin_matches_oracle_fp64():
    """The IM proxy of every netting set of a book (each valued on every date of the book's
    simulation grid) against the oracle's book restatement."""
    from zlatko_amd import native
    spec = W.book_case(num_paths=900)
    e = native.Engine()
    e.set_book(True)
    e.set_initial_margin(0.99, 10.0 / 250.0)
    plan = e.compile(spec)
    plan.execute()
    for s in range(len(spec.sets)):
        got = plan.initial_margin(s)
        want = oracle_lib.oracle_book_initial_margin(spec, s, 0.99)
        assert len(got) == len(want) and got[0] == 0.0
        err = np.abs(got - want).max() / max(np.abs(want).max(), 1.0)
        print("book IM", s, f"{err:.2e}", flush=True)
        assert err <= FP64_TOL
    plan.close()
    e.close()


def test_graph_mode_replays_match_eager_executes():
    """Graph mode: each plan's execute captured once as a HIP graph and replayed. Three replays of
    each plan (a Bermudan/AMC plan, an FP32 CSA plan, per-trade and IM outputs), with eager
    executes of the other plans and synchronisations between them, give the eager profiles bit
    for bit; a plan whose shared workspace grew is captured again."""
    from zlatko_amd import native
    specs = [W.small_reference_case(num_paths=600, csa=True, fx=True, credits=True, bermudans=3),
             W.c2(num_paths=700, num_sets=3, num_trades=30, weeks=52, horizon=1.0)]
    eager = native.Engine()
    eager.set_initial_margin(0.99)
    want = [eager.compile(specs[0]).execute(), eager.compile(specs[1], abi.CCR_FP32).execute()]
    e = native.Engine()
    e.set_initial_margin(0.99)
    e.set_shared_workspace(True)
    plans = [e.compile(specs[0]), e.compile(specs[1], abi.CCR_FP32)]
    e.set_graph(True)
    for rep in range(3):
        for i, p in enumerate(plans):
            got = p.execute()
            for o, w in zip(got, want[i]):
                for k in abi.PROFILE_FIELDS:
                    assert np.array_equal(o.get(k), w.get(k)), (rep, i, k)
            assert p.initial_margin(0)[0] == 0.0
        e.set_graph(False)
        plans[rep % 2].execute()  # eager work between replays
        e.synchronize()
        e.set_graph(True)
    bigger = e.compile(W.small_reference_case(num_paths=2000, csa=True, bermudans=2))
    bigger.execute()  # grows the shared workspace: the next replay recaptures
    for o, w in zip(plans[0].execute(), want[0]):
        for k in abi.PROFILE_FIELDS:
            assert np.array_equal(o.get(k), w.get(k)), k
    for p in plans + [bigger]:
        p.close()
    e.close()
    eager.close()
