"""Nested conditional initial margin on the GPU against the FP64 oracle (same inner generator,
same outer stream): E[D IM], E[IM] and the 5 / 50 / 95% quantiles of IM per reporting date."""
import numpy as np
import pytest

from tests import oracle_lib as OL
from zlatko_amd import abi, native
from zlatko_amd import workloads as W

pytestmark = pytest.mark.gpu

MPOR = 10.0 / 250.0
FIELDS = ("discounted", "mean", "band05", "band50", "band95")


def _windowed_spec(num_paths=512):
    """Two-leg swaps (quarterly floating legs from 0), FX forwards and options, with reporting dates
    just before floating-period starts, so that margin-period windows hold fixing dates."""
    spec = W.small_reference_case(num_paths=num_paths, dates=10, fx=True, options=2, two_leg=True)
    extra = np.array([0.99, 1.49, 2.24, 3.0 - 0.02])
    return abi.Spec(spec.model, spec.sets, np.concatenate([spec.base_grid, extra]), num_paths, 7, True, 0.95, "win")


def _errors(got, ref):
    return {k: float(np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300)) for k in FIELDS}


def _run(spec, inner, q, precision=abi.CCR_FP64):
    e = native.Engine()
    try:
        e.set_nested_initial_margin(inner, q, MPOR)
        plan = e.compile(spec, precision)
        plan.execute()
        out = [plan.nested_initial_margin(s) for s in range(len(spec.sets))]
        plan.close()
        return out
    finally:
        e.close()


@pytest.mark.parametrize("inner,q", [(64, 0.9), (256, 0.99)])
def test_windowed_case_matches_oracle_fp64(inner, q):
    spec = _windowed_spec()
    got = _run(spec, inner, q)[0]
    ref = OL.oracle_nested_im(spec, 0, inner, q, MPOR, threads=8)
    errs = _errors(got, ref)
    print("nested im fp64", inner, q, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert len(got["mean"]) == len(ref["mean"])
    assert max(errs.values()) <= 1e-10
    # No window past the horizon minus the margin period (nor after the last cash flow).
    assert got["mean"][-1] == 0.0 and (got["mean"] > 0.0).sum() >= len(got["mean"]) // 2


def test_three_currency_sets_match_oracle_fp64():
    """Foreign rates, FX and a weekly grid (windows of two steps); two netting sets in one plan."""
    spec = W.c2(num_paths=256, weeks=52, num_sets=2, num_trades=20)
    outs = _run(spec, 64, 0.9)
    for s in range(2):
        ref = OL.oracle_nested_im(spec, s, 64, 0.9, MPOR, threads=8)
        errs = _errors(outs[s], ref)
        print("nested im 3-ccy set", s, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert max(errs.values()) <= 1e-10


def test_fp32_plan_error_vs_oracle():
    spec = _windowed_spec()
    got = _run(spec, 64, 0.9, abi.CCR_FP32)[0]
    ref = OL.oracle_nested_im(spec, 0, 64, 0.9, MPOR, threads=8)
    errs = _errors(got, ref)
    print("nested im fp32", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 2e-3


def test_nested_im_errors():
    e = native.Engine()
    try:
        for inner, q in ((1, 0.99), (2000, 0.99), (64, 0.5), (64, 1.5)):
            with pytest.raises(native.CcrError):
                e.set_nested_initial_margin(inner, q, MPOR)
        with pytest.raises(native.CcrError):
            e.set_nested_initial_margin(64, 0.9, 0.0)
        e.set_nested_initial_margin(64, 0.9, MPOR)
        with pytest.raises(native.CcrError) as ex:
            e.compile(W.small_reference_case(num_paths=64, dates=6, bermudans=1))
        assert ex.value.code == abi.CCR_ERR_UNSUPPORTED
        e.set_book(True)
        with pytest.raises(native.CcrError) as ex:
            e.compile(W.book_case(num_paths=64))
        assert ex.value.code == abi.CCR_ERR_UNSUPPORTED
        e.set_book(False)
        e.set_nested_initial_margin(0, 0.9, MPOR)
        plan = e.compile(W.small_reference_case(num_paths=64, dates=6))
        plan.execute()
        with pytest.raises(native.CcrError):
            plan.nested_initial_margin(0)
        plan.close()
    finally:
        e.close()
