"""The FP64 restatement of the reference's MLP surrogate training (oracle/mlp_oracle.cpp), pinned
by the reference's own fitMlp (golden fixture and, where built, the live reference), by
Standardiser semantics and, for the SiLU / deeper extension, by central finite differences."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as OL
from zlatko_amd import abi

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def test_fit_mlp_matches_reference_golden_bit_for_bit():
    g = GOLDEN["mlp"]
    X, Y = OL.mlp_dataset(g["n"], g["din"], g["dout"])
    d = abi.MlpDescC(g["din"], g["hidden"], 2, g["dout"], abi.CCR_ACT_TANH)
    p, s = OL.oracle_fit_mlp(d, X, Y, g["epochs"], g["lr"], g["wd"], g["seed"])
    assert np.array_equal(p, np.array(g["params"]))
    assert np.array_equal(s, np.array(g["stdz"]))


@pytest.mark.skipif("OL.ref() is None", reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("hidden,epochs,seed", [(8, 1, 3), (16, 40, 7), (12, 120, 11)])
def test_fit_mlp_matches_live_reference(hidden, epochs, seed):
    X, Y = OL.mlp_dataset(70, 4, 3, seed=seed)
    d = abi.MlpDescC(4, hidden, 2, 3, abi.CCR_ACT_TANH)
    p, s = OL.oracle_fit_mlp(d, X, Y, epochs, 0.02, 1e-5, seed)
    pr, sr = np.zeros_like(p), np.zeros_like(s)
    assert OL.ref().ccr_ref_fit_mlp(abi.dptr(X), abi.dptr(Y), 70, 4, 3, hidden, epochs, 0.02, 1e-5, seed,
                                    abi.dptr(pr), abi.dptr(sr)) == 0
    assert np.array_equal(p, pr) and np.array_equal(s, sr)


def test_standardiser_population_sd_and_constant_columns():
    rows = np.array([[1.0, 5.0], [3.0, 5.0], [5.0, 5.0], [7.0, 5.0]])
    m, sd = np.zeros(2), np.zeros(2)
    OL.oracle().ccr_oracle_standardiser_fit(abi.dptr(rows), 4, 2, abi.dptr(m), abi.dptr(sd))
    assert m[0] == 4.0 and m[1] == 5.0
    assert abs(sd[0] - np.sqrt(5.0)) < 1e-15 and sd[1] == 1.0


@pytest.mark.parametrize("act,layers", [(abi.CCR_ACT_TANH, 2), (abi.CCR_ACT_SILU, 3), (abi.CCR_ACT_SILU, 1)])
def test_gradient_matches_central_differences(act, layers):
    rng = np.random.default_rng(5)
    d = abi.MlpDescC(3, 7, layers, 2, act)
    n = abi.mlp_param_count(d)
    p = np.zeros(n)
    OL.oracle().ccr_oracle_mlp_init(C.byref(d), 4, abi.dptr(p))
    p += 0.05 * rng.standard_normal(n)  # non-zero biases too
    xs = rng.standard_normal((37, 3))
    ys = rng.standard_normal((37, 2))
    loss, g = OL.oracle_mlp_gradient(d, p, xs, ys)
    h = 1e-6
    for q in rng.choice(n, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        lp, _ = OL.oracle_mlp_gradient(d, pp, xs, ys)
        lm, _ = OL.oracle_mlp_gradient(d, pm, xs, ys)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])), (q, fd, g[q])
    assert loss > 0


def test_training_reduces_loss_and_predict_inverts_standardisation():
    X, Y = OL.mlp_dataset(200, 4, 2, seed=3)
    d = abi.MlpDescC(4, 32, 3, 2, abi.CCR_ACT_SILU)
    p0, s0 = OL.oracle_fit_mlp(d, X, Y, 0)
    p1, s1 = OL.oracle_fit_mlp(d, X, Y, 300)
    xs = (X - s1[:4]) / s1[4:8]
    ys = (Y - s1[8:10]) / s1[10:12]
    l0, _ = OL.oracle_mlp_gradient(d, p0, xs, ys)
    l1, _ = OL.oracle_mlp_gradient(d, p1, xs, ys)
    assert l1 < 0.1 * l0
    pred = np.zeros_like(Y)
    OL.oracle().ccr_oracle_mlp_predict(C.byref(d), abi.dptr(p1), abi.dptr(s1), abi.dptr(X), 200, abi.dptr(pred))
    assert np.sqrt(np.mean((pred - Y) ** 2)) < 0.5 * np.std(Y)
