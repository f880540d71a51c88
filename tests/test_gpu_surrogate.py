"""The exposure surrogate: an MLP trained on the GPU on the values the regression method produced,
gated on hold-out paths, with the regression method as fallback."""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_lib as OL
from zlatko_amd import abi
from zlatko_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hidden_layers,outputs", [(3, 1), (4, 1), (5, 1), (3, 40), (3, 128)])
def test_bf16_fused_inference_matches_fp64_predict(engine, hidden_layers, outputs):
    """The fused inference kernel (weights in LDS) against the FP64 restatement of Mlp::predict, at
    every depth the BF16 path trains (4 and 5 hidden layers leave room in LDS for fewer wavefronts
    per workgroup than 3 do, and for one sample tile per wavefront at 5) and at output widths up to
    128 (one output per netting set of a book)."""
    desc = abi.MlpDescC(7, 128, hidden_layers, outputs, abi.CCR_ACT_SILU)
    X, Y = OL.mlp_dataset(3001, 7, outputs, seed=4)
    p, s = OL.oracle_fit_mlp(desc, X, Y, 20, lr=0.003)
    ref = np.zeros_like(Y)
    OL.oracle().ccr_oracle_mlp_predict(C.byref(desc), abi.dptr(p), abi.dptr(s), abi.dptr(X), X.shape[0], abi.dptr(ref))
    got = engine.mlp_predict(desc, abi.CCR_BF16, p, s, X)
    # BF16 weights and activations: errors of a few 1e-3 of the output scale per layer; the worst
    # of 3,001 x outputs values grows with the count, the RMS does not. A misplaced output or
    # sample is off by the output's whole spread.
    err = (np.abs(got - ref).max(axis=0) / Y.std(axis=0)).max()
    rms = np.sqrt(np.mean((got - ref) ** 2)) / Y.std()
    print("fused bf16 inference", hidden_layers, outputs, f"max {err:.2e} rms {rms:.2e}", flush=True)
    assert err < (3e-2 if outputs == 1 else 6e-2) and rms < 1e-2


def test_gate_rejects_and_falls_back_bit_exact(engine):
    """With a zero tolerance the gate keeps the regression method: identical results."""
    tr = engine.compile(W.c4(num_paths=3000, dates=40, seed=101))
    ho = engine.compile(W.c4(num_paths=3000, dates=40, seed=202))
    sur = engine.surrogate_train(tr, ho, train_paths=1000, epochs=5, lr=0.001, batch=4096, gate=0.0)
    assert sur.report.candidate_sets == 1 and sur.report.enabled_sets == 0
    enabled, err = sur.set_info(0)
    assert enabled == 0 and err > 0.0
    prod = engine.compile(W.c4(num_paths=3000, dates=40, seed=42))
    a = prod.execute()[0].as_dict()
    b = prod.execute_surrogate(sur)[0].as_dict()
    for k in abi.PROFILE_FIELDS:
        assert np.array_equal(a[k], b[k]), k
    for p in (tr, ho, prod):
        p.close()
    sur.close()


@pytest.mark.parametrize("hidden_layers", [4, 5])
def test_deeper_surrogates_train_and_execute(engine, hidden_layers):
    """Every depth the BF16 path accepts for training also runs the surrogate's fused inference:
    with the gate open the surrogate execute succeeds and the set keeps the surrogate; its EE stays
    within the gate's hold-out error plus Monte Carlo noise of the plain execute."""
    tr = engine.compile(W.c4(num_paths=3000, dates=40, seed=101))
    ho = engine.compile(W.c4(num_paths=3000, dates=40, seed=202))
    sur = engine.surrogate_train(tr, ho, train_paths=1000, hidden_layers=hidden_layers, epochs=20, lr=0.001,
                                 batch=4096, gate=1.0)
    assert sur.report.enabled_sets == 1
    prod = engine.compile(W.c4(num_paths=3000, dates=40, seed=42))
    a = prod.execute()[0]
    b = prod.execute_surrogate(sur)[0]
    ee_a, ee_b = a.get("expected_exposure"), b.get("expected_exposure")
    err = np.abs(ee_b - ee_a).max() / ee_a.max()
    print("surrogate depth", hidden_layers, f"hold-out {sur.report.max_ee_error:.3e} production {err:.3e}", flush=True)
    assert np.all(np.isfinite(ee_b)) and err < sur.report.max_ee_error + 0.05
    for p in (tr, ho, prod):
        p.close()
    sur.close()


def test_surrogate_fp32_cva_against_plain_fp64(engine):
    """Config 4 shape (two-leg swaps, FX forwards and Bermudans, CIR++ credit with wrong-way
    correlation): the FP32 path with the surrogate valuing the Bermudans against the plain FP64
    method on the same production paths. States the EE and CVA errors."""
    tr = engine.compile(W.c4(num_paths=20000, seed=101))
    ho = engine.compile(W.c4(num_paths=20000, seed=202))
    sur = engine.surrogate_train(tr, ho, train_paths=2000, epochs=80, lr=0.0005, batch=4096, gate=0.02)
    rep = sur.report
    print("surrogate", f"samples {rep.samples} train {rep.train_ms:.0f} ms loss {rep.loss:.3e}",
          f"hold-out EE error {rep.max_ee_error:.4f} enabled {rep.enabled_sets}/{rep.candidate_sets}", flush=True)
    assert rep.samples >= 200000 and rep.enabled_sets == 1
    spec = W.c4(num_paths=20000, seed=42)
    p64 = engine.compile(spec, abi.CCR_FP64)
    p32 = engine.compile(spec, abi.CCR_FP32)
    ref = p64.execute()[0]
    out = p32.execute_surrogate(sur)[0]
    ee_err = np.abs(out.get("expected_exposure") - ref.get("expected_exposure")).max() / ref.get("expected_exposure").max()
    cva_ref = ref.get("cva_increments")[1:].sum()
    cva = out.get("cva_increments")[1:].sum()
    print("fp32+surrogate vs plain fp64", f"EE error {ee_err:.3e}", f"CVA {cva:.6e} vs {cva_ref:.6e}",
          f"rel {abs(cva - cva_ref) / cva_ref:.3e}", flush=True)
    assert ee_err < 0.02 and abs(cva - cva_ref) < 0.005 * cva_ref
    for p in (tr, ho, p64, p32):
        p.close()
    sur.close()


def _book(num_paths, seed, num_sets=8, largest=40, dates=40):
    return W.c5(num_paths=num_paths, num_sets=num_sets, largest=largest, dates=dates, seed=seed, bermudan_share=0.25)


def _berm_sets(spec):
    return [i for i, s in enumerate(spec.sets) if any(t.kind == abi.CCR_TRADE_BERMUDAN_SWAPTION for t in s.trades)]


def test_book_surrogate_one_network_gate_per_set(engine):
    """A portfolio book (sets on the book's grids, one simulated state): one network whose outputs
    are the Bermudan values of every set holding Bermudans. With the gate closed every set keeps
    the regression method (bit-identical results); with it open every set takes the network, each
    set's EE stays within its hold-out error plus Monte Carlo noise of the plain execute, and sets
    without Bermudans are untouched (bit-identical)."""
    engine.set_book(True)
    try:
        tr, ho = engine.compile(_book(4000, 101)), engine.compile(_book(4000, 202))
        prod_spec = _book(4000, 42)
        berm = _berm_sets(prod_spec)
        assert 2 <= len(berm) < len(prod_spec.sets)
        prod = engine.compile(prod_spec)
        plain = [o.as_dict() for o in prod.execute()]
        closed = engine.surrogate_train(tr, ho, train_paths=1000, epochs=5, lr=0.001, batch=4096, gate=0.0)
        assert closed.report.networks == 1 and closed.report.candidate_sets == len(berm)
        assert closed.report.enabled_sets == 0
        for a, b in zip(plain, prod.execute_surrogate(closed)):
            for k in abi.PROFILE_FIELDS:
                assert np.array_equal(a[k], b.as_dict()[k]), k
        closed.close()
        sur = engine.surrogate_train(tr, ho, train_paths=2000, epochs=40, lr=0.0005, batch=4096, gate=1.0)
        rep = sur.report
        print("book surrogate", f"networks {rep.networks} outputs {rep.candidate_sets} samples {rep.samples}",
              f"train {rep.train_ms:.0f} ms hold-out {rep.max_ee_error:.3e}", flush=True)
        assert rep.networks == 1 and rep.enabled_sets == len(berm)
        got = [o.as_dict() for o in prod.execute_surrogate(sur)]
        for i, (a, b) in enumerate(zip(plain, got)):
            enabled, herr = sur.set_info(i)
            if i not in berm:
                assert enabled == -1
                for k in abi.PROFILE_FIELDS:
                    assert np.array_equal(a[k], b[k]), (i, k)
                continue
            assert enabled == 1
            ee_a, ee_b = a["expected_exposure"], b["expected_exposure"]
            err = np.abs(ee_b - ee_a).max() / np.abs(ee_a).max()
            print("set", i, f"hold-out {herr:.3e} production {err:.3e}", flush=True)
            assert np.all(np.isfinite(ee_b)) and err < herr + 0.05
        sur.close()
        for p in (tr, ho, prod):
            p.close()
    finally:
        engine.set_book(False)


def test_book_surrogate_stage_against_regressions(engine):
    """What the learned part gains over the method on a 50-set portfolio book (50,000 paths x 120
    dates, FP32 plans): the time of the Bermudan stage (states pass plus the regressions, or plus
    the network) of one execute with the surrogate against the same execute with the American
    Monte Carlo regressions, and the CVA of the FP32 + surrogate run against the plain FP64 run.
    States both; asserts only that the surrogate stage ran for every set (a timing ratio is a
    measurement, reported)."""
    engine.set_book(True)
    try:
        mk = lambda p, seed: W.c5(num_paths=p, num_sets=50, seed=seed)  # noqa: E731
        tr, ho = engine.compile(mk(8000, 101)), engine.compile(mk(8000, 202))
        sur = engine.surrogate_train(tr, ho, train_paths=4000, epochs=60, lr=0.0005, batch=4096, gate=1.0)
        rep = sur.report
        spec = mk(50000, 42)
        p64, p32 = engine.compile(spec, abi.CCR_FP64), engine.compile(spec, abi.CCR_FP32)

        def stage(run):
            run()
            engine.synchronize()
            t = []
            for _ in range(3):
                run()
                t.append(engine.timing())
            return min(x.amc_ms for x in t), min(x.total_ms for x in t)

        amc32, tot32 = stage(p32.execute_async)
        sur32, stot32 = stage(lambda: p32.execute_surrogate_async(sur))
        _, tot64 = stage(p64.execute_async)
        ref = [o.get("cva_increments")[1:].sum() for o in p64.execute()]
        got = [o.get("cva_increments")[1:].sum() for o in p32.execute_surrogate(sur)]
        cva_err = abs(sum(got) - sum(ref)) / abs(sum(ref))
        print("book surrogate", f"networks {rep.networks} outputs {rep.candidate_sets} hold-out {rep.max_ee_error:.3e}",
              f"train {rep.train_ms:.0f} ms", flush=True)
        print(f"Bermudan stage: regressions {amc32:.2f} ms, surrogate {sur32:.2f} ms ({amc32 / sur32:.2f}x); "
              f"execute: FP64 plain {tot64:.2f} ms, FP32 {tot32:.2f} ms, FP32 + surrogate {stot32:.2f} ms "
              f"({tot64 / stot32:.2f}x vs FP64); book CVA error {cva_err:.3e}", flush=True)
        assert rep.enabled_sets == rep.candidate_sets and rep.networks == 1
        assert np.isfinite(cva_err)
        for p in (tr, ho, p64, p32):
            p.close()
        sur.close()
    finally:
        engine.set_book(False)
