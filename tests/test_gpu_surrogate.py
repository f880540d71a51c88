"""The exposure surrogate: an MLP trained on the GPU on the values the regression method produced,
gated on hold-out paths, with the regression method as fallback."""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_lib as OL
from zlatko_amd import abi
from zlatko_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hidden_layers", [3, 4, 5])
def test_bf16_fused_inference_matches_fp64_predict(engine, hidden_layers):
    """The fused inference kernel (weights in LDS) against the FP64 restatement of Mlp::predict, at
    every depth the BF16 path trains (4 and 5 hidden layers leave room in LDS for fewer wavefronts
    per workgroup than 3 do)."""
    desc = abi.MlpDescC(7, 128, hidden_layers, 1, abi.CCR_ACT_SILU)
    X, Y = OL.mlp_dataset(3001, 7, 1, seed=4)
    p, s = OL.oracle_fit_mlp(desc, X, Y, 20, lr=0.003)
    ref = np.zeros_like(Y)
    OL.oracle().ccr_oracle_mlp_predict(C.byref(desc), abi.dptr(p), abi.dptr(s), abi.dptr(X), X.shape[0], abi.dptr(ref))
    got = engine.mlp_predict(desc, abi.CCR_BF16, p, s, X)
    err = np.abs(got - ref).max() / Y.std()
    print("fused bf16 inference", hidden_layers, f"{err:.2e}", flush=True)
    assert err < 3e-2


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
