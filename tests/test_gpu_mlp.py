"""The GPU MLP surrogate kernels against the FP64 restatement of the reference's fitMlp.

FP32 path: gradients within 1e-5 (scale-relative) of the FP64 oracle on the reference
architecture (5-16-16-6 tanh) and on 3 x 128 SiLU; training follows the reference's golden fit.
BF16 path (matrix cores, FP32 accumulation): gradients within 3e-2 (norm-relative) of the FP64
oracle; training converges.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as OL
from zlatko_amd import abi

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def _standardised(X, Y):
    xs = (X - X.mean(0)) / X.std(0)
    ys = (Y - Y.mean(0)) / Y.std(0)
    return np.ascontiguousarray(xs), np.ascontiguousarray(ys)


def _init(desc, seed):
    p = np.zeros(abi.mlp_param_count(desc))
    OL.oracle().ccr_oracle_mlp_init(C.byref(desc), seed, abi.dptr(p))
    return p


@pytest.mark.parametrize("shape", [(5, 16, 2, 6, abi.CCR_ACT_TANH), (4, 128, 3, 2, abi.CCR_ACT_SILU),
                                   (3, 40, 1, 5, abi.CCR_ACT_SILU)])
def test_fp32_gradient_matches_fp64_oracle(engine, shape):
    desc = abi.MlpDescC(*shape)
    X, Y = OL.mlp_dataset(1000, desc.inputs, desc.outputs, seed=13)
    xs, ys = _standardised(X, Y)
    p = _init(desc, 3)
    p += 0.01 * np.random.default_rng(1).standard_normal(p.size)
    lo, go = OL.oracle_mlp_gradient(desc, p, xs, ys)
    lg, gg = engine.mlp_gradient(desc, abi.CCR_FP32, p, xs, ys)
    err = np.abs(gg - go).max() / np.abs(go).max()
    print("fp32 grad", shape, f"{err:.2e}", f"loss {lg:.6e} vs {lo:.6e}", flush=True)
    assert err < 1e-5 and abs(lg - lo) < 1e-5 * lo


@pytest.mark.parametrize("shape,n", [((4, 128, 3, 2, abi.CCR_ACT_SILU), 3000), ((8, 128, 2, 3, abi.CCR_ACT_TANH), 777),
                                     ((30, 128, 1, 32, abi.CCR_ACT_SILU), 64)])
def test_bf16_gradient_matches_fp64_oracle(engine, shape, n):
    desc = abi.MlpDescC(*shape)
    X, Y = OL.mlp_dataset(n, desc.inputs, desc.outputs, seed=17)
    xs, ys = _standardised(X, Y)
    p = _init(desc, 5)
    lo, go = OL.oracle_mlp_gradient(desc, p, xs, ys)
    lg, gg = engine.mlp_gradient(desc, abi.CCR_BF16, p, xs, ys)
    err = np.linalg.norm(gg - go) / np.linalg.norm(go)
    print("bf16 grad", shape, n, f"{err:.2e}", f"loss {lg:.6e} vs {lo:.6e}", flush=True)
    assert err < 3e-2 and abs(lg - lo) < 2e-2 * lo


def test_fp32_training_follows_the_reference_fit():
    """Reference architecture, reference dataset and hyper-parameters (golden fixture of the
    reference's own fitMlp): FP32 training lands on the same fit (predictions within 1e-3 of the
    data scale) with a loss within 1% of the reference's."""
    from zlatko_amd import native
    g = GOLDEN["mlp"]
    X, Y = OL.mlp_dataset(g["n"], g["din"], g["dout"])
    desc = abi.MlpDescC(g["din"], g["hidden"], 2, g["dout"], abi.CCR_ACT_TANH)
    e = native.Engine()
    p, s, loss, _ = e.mlp_train(desc, X, Y, g["epochs"], g["lr"], g["wd"], g["seed"], abi.CCR_FP32)
    ref_p, ref_s = np.array(g["params"]), np.array(g["stdz"])
    assert np.array_equal(s, ref_s)  # standardisers are fitted on the host in FP64
    pred = e.mlp_predict(desc, abi.CCR_FP32, p, s, X)
    ref_pred = np.zeros_like(Y)
    OL.oracle().ccr_oracle_mlp_predict(C.byref(desc), abi.dptr(ref_p), abi.dptr(ref_s), abi.dptr(X), X.shape[0],
                                       abi.dptr(ref_pred))
    xs, ys = (X - s[:5]) / s[5:10], (Y - s[10:16]) / s[16:22]
    ref_loss, _ = OL.oracle_mlp_gradient(desc, ref_p, np.ascontiguousarray(xs), np.ascontiguousarray(ys))
    dev = np.abs(pred - ref_pred).max() / np.abs(Y).max()
    print("fp32 fit", f"loss {loss:.6e} ref {ref_loss:.6e}", f"pred dev {dev:.2e}", flush=True)
    e.close()
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss and dev < 1e-3


def test_bf16_training_converges_and_predicts(engine):
    desc = abi.MlpDescC(4, 128, 3, 2, abi.CCR_ACT_SILU)
    X, Y = OL.mlp_dataset(20000, 4, 2, seed=21)
    p0, s0, l0, _ = engine.mlp_train(desc, X, Y, 0, precision=abi.CCR_BF16)
    p, s, loss, ms = engine.mlp_train(desc, X, Y, 400, lr=0.003, precision=abi.CCR_BF16)
    pred = engine.mlp_predict(desc, abi.CCR_BF16, p, s, X)
    rmse = np.sqrt(np.mean((pred - Y) ** 2)) / Y.std()
    print("bf16 fit", f"loss {l0:.4e} -> {loss:.4e}", f"rmse/sd {rmse:.3e}", f"{ms:.1f} ms", flush=True)
    assert loss < 0.05 * l0 and rmse < 0.25


def test_predict_fp32_matches_oracle(engine):
    desc = abi.MlpDescC(6, 32, 2, 3, abi.CCR_ACT_SILU)
    X, Y = OL.mlp_dataset(500, 6, 3, seed=2)
    p, s = OL.oracle_fit_mlp(desc, X, Y, 50)
    ref = np.zeros_like(Y)
    OL.oracle().ccr_oracle_mlp_predict(C.byref(desc), abi.dptr(p), abi.dptr(s), abi.dptr(X), 500, abi.dptr(ref))
    got = engine.mlp_predict(desc, abi.CCR_FP32, p, s, X)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
