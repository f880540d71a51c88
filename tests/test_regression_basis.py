"""The swap-rate / annuity regression basis (an extension: the reference regresses on the state
alone, src/instruments/bermudan_swaption.cpp:77-93) in the oracle, pinned without retyping it:

* the fit is ridge least squares -- solved here by QR on the augmented system [Phi; sqrt(ridge)],
  a different route from the oracle's normal equations and Cholesky;
* on independent regressors it recovers a cubic in (s, a) term by term (a dropped, swapped or
  mis-signed monomial fails);
* its degree rules (sample size, degenerate regressor, empty sample);
* the regressors: a payer swap struck at the swap rate minus d is worth N d A at the valuation
  date, through the oracle's swap valuation that the reference's golden vectors pin;

and the product's plan records and basis arithmetic (run on the CPU by the plan emulator) against
the oracle with either basis."""
import ctypes as C

import numpy as np
import pytest

from tests import compare, oracle_lib
from tests.test_plan_emulator import emu, run_emu  # noqa: F401  (module fixture)
from zlatko_amd import abi
from zlatko_amd import workloads as W

RIDGE = 1e-4


def _sa_fit(u1, u2, y, mask=None, degree=3):
    n = len(u1)
    coef = np.zeros(10)
    params = np.zeros(5)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    rc = oracle_lib.oracle().ccr_oracle_sa_regression(
        abi.dptr(np.ascontiguousarray(u1, np.float64)), abi.dptr(np.ascontiguousarray(u2, np.float64)),
        abi.dptr(np.ascontiguousarray(y, np.float64)), None if m is None else m.ctypes.data_as(C.c_void_p), n,
        degree, coef.ctypes.data_as(C.c_void_p), params.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return coef, params


def _monomials(s, a):
    return np.stack([np.ones_like(s), s, a, s * s, s * a, a * a, s ** 3, s * s * a, s * a * a, a ** 3], 1)


def test_sa_regression_is_ridge_least_squares():
    rng = np.random.default_rng(5)
    n = 6000
    u1 = rng.normal(0.03, 0.004, n)
    u2 = 4.0 + 0.5 * rng.normal(size=n) + 30.0 * (u1 - 0.03)  # correlated, not degenerate
    y = 1e6 * np.maximum(u1 - 0.03, 0) * u2 + 1e3 * rng.normal(size=n)
    mask = rng.random(n) < 0.7
    coef, params = _sa_fit(u1, u2, y, mask)
    x1, x2, yy = u1[mask], u2[mask], y[mask]
    m1, m2 = x1.mean(), x2.mean()
    s1, s2 = x1.std(), x2.std()
    assert np.allclose(params, [m1, s1, m2, s2, 3], rtol=1e-12)
    phi = _monomials((x1 - m1) / s1, (x2 - m2) / s2)
    g = np.einsum("pi,pi->i", phi, phi)
    ridge = 1e-12 * len(yy) + RIDGE * g
    aug = np.vstack([phi, np.diag(np.sqrt(ridge))])
    want = np.linalg.lstsq(aug, np.concatenate([yy, np.zeros(10)]), rcond=None)[0]
    assert np.abs(coef - want).max() <= 1e-9 * np.abs(want).max(), (coef, want)


def test_sa_regression_recovers_a_cubic_term_by_term():
    rng = np.random.default_rng(6)
    n = 20000
    u1 = rng.uniform(-1.0, 1.0, n) * 0.01 + 0.03
    u2 = rng.uniform(2.0, 6.0, n)
    m1, s1, m2, s2 = u1.mean(), u1.std(), u2.mean(), u2.std()
    truth = np.array([1.0, -2.0, 3.0, 0.5, -1.5, 0.25, 0.75, -0.6, 0.4, -0.2])
    y = _monomials((u1 - m1) / s1, (u2 - m2) / s2) @ truth
    coef, params = _sa_fit(u1, u2, y)
    assert params[4] == 3
    # The ridge of 1e-4 x the diagonal biases the coefficients by a few 1e-3 here (a dropped or swapped term
    # would be off by its whole coefficient, >= 0.2).
    assert np.abs(coef - truth).max() <= 1e-2, coef - truth


def test_sa_regression_degree_rules():
    rng = np.random.default_rng(7)
    u1, u2, y = rng.normal(size=40), rng.normal(size=40), rng.normal(size=40)
    # 12 samples: 10 terms need 20, 6 terms need 12 -> total degree 2.
    coef, params = _sa_fit(u1[:12], u2[:12], y[:12])
    assert params[4] == 2 and np.all(coef[6:] == 0.0)
    # 5 samples: even degree 1 (3 terms) needs 6 -> the sample mean.
    coef, params = _sa_fit(u1[:5], u2[:5], y[:5])
    assert params[4] == 0 and coef[0] == pytest.approx(y[:5].mean(), rel=1e-15) and np.all(coef[1:] == 0.0)
    # A degenerate regressor drops the fit to the mean.
    coef, params = _sa_fit(u1, np.full(40, 3.0), y)
    assert params[4] == 0 and coef[0] == pytest.approx(y.mean(), rel=1e-15)
    # ... also when its one-pass variance is rounding noise (a constant swap rate on a date where
    # every path has the same state), or when it varies by 1e-9 of its size.
    coef, params = _sa_fit(np.full(40, 0.1) * (1.0 + 1e-15 * rng.normal(size=40)), u2, y)
    assert params[4] == 0
    coef, params = _sa_fit(np.full(40, 0.1) * (1.0 + 1e-9 * rng.normal(size=40)), u2, y)
    assert params[4] == 0
    coef, params = _sa_fit(np.full(40, 0.1) * (1.0 + 1e-4 * rng.normal(size=40)), u2, y)
    assert params[4] == 3
    # The trade's degree caps the total degree; an empty sample predicts 0.
    coef, params = _sa_fit(u1, u2, y, degree=1)
    assert params[4] == 1 and np.all(coef[3:] == 0.0)
    coef, params = _sa_fit(u1, u2, y, mask=np.zeros(40, bool))
    assert params[4] == 0 and np.all(coef == 0.0)


@pytest.mark.parametrize("delta", [0.0, 0.0025])
def test_swap_rate_and_annuity_price_a_swap(delta):
    """A payer swap struck at S - d is worth N d A today (x = 0): the regressors of the basis
    against the oracle's swap valuation (pinned to the reference's golden vectors)."""
    curve = W.Curve([0.5, 2.0, 5.0, 10.0], [0.020, 0.025, 0.030, 0.033])
    hw = W.HullWhite(curve, 0.05, 0.01)
    sched = np.array([1.5, 2.5, 3.5, 4.5, 5.5, 6.5])
    spec0 = W.Spec(W.Model(hw), [W.NettingSet([W.Trade(abi.CCR_TRADE_SWAP, 1e7, 0.03, True, schedule=sched)])],
                   W.uniform_grid(6.5, 4), 8, 1, True, 0.95, "swap")
    S, A = C.c_double(), C.c_double()
    rc = oracle_lib.oracle().ccr_oracle_swap_rate_annuity(C.byref(spec0.generator.rates), abi.dptr(sched), len(sched),
                                                           0.0, 0.0, C.byref(S), C.byref(A))
    assert rc == 0 and 0.01 < S.value < 0.06 and 3.0 < A.value < 5.0
    N = 1e7
    spec = W.Spec(W.Model(hw), [W.NettingSet([W.Trade(abi.CCR_TRADE_SWAP, N, S.value - delta, True,
                                                      schedule=sched)])],
                  W.uniform_grid(6.5, 4), 8, 1, True, 0.95, "swap")
    out = oracle_lib.run_oracle(spec)[0]
    assert out.get("times")[0] == 0.0
    v0 = out.get("expected_value")[0]
    assert v0 == pytest.approx(N * delta * A.value, rel=1e-12, abs=1e-7 * N)


def _set_basis(emu_lib, basis):
    emu_lib.emu_set_regression_basis.argtypes = [C.c_int32]
    emu_lib.emu_set_regression_basis(basis)
    oracle_lib.oracle().ccr_oracle_set_regression_basis(basis)


CASES = {
    "c3_small": lambda: W.c3(num_paths=96, dates=30, num_swaptions=12),
    "c3_full_shape": lambda: W.c3(num_paths=24),
    "csa_mix": lambda: W.small_reference_case(120, csa=True, fx=True, bermudans=3),
    "cash": lambda: W.small_reference_case(64, bermudans=3, cash_settled=True),
    "tiny": lambda: W.small_reference_case(5, csa=True, bermudans=2),
}


@pytest.mark.parametrize("basis", [abi.CCR_BASIS_STATE, abi.CCR_BASIS_SWAP_RATE_ANNUITY])
@pytest.mark.parametrize("case", sorted(CASES))
def test_emulated_engine_matches_oracle(emu, case, basis):  # noqa: F811
    """Regression records (exercise k of every swaption, the continuation of every valued date),
    the tails' annuity / float-leg terms and the basis arithmetic, on the CPU in path order,
    against the oracle with the same basis. The state basis reproduces the reference's fits to
    rounding. The swap-rate / annuity regressors come from the plan's bond polynomials here and
    from exact zero bonds in the oracle (~1e-14 apart); the ridge of 1e-4 x the diagonal bounds the
    condition of the normal equations near 1e6, so fits and profiles agree to ~1e-10 (bar 1e-8)."""
    spec = CASES[case]()
    try:
        _set_basis(emu, basis)
        e = run_emu(emu, spec)
        o = oracle_lib.run_oracle(spec, threads=4, rows=True)
    finally:
        _set_basis(emu, -1)
    tol = 1e-12 if basis == abi.CCR_BASIS_STATE else 1e-8
    for s in range(len(spec.sets)):
        errs = compare.field_errors(e[s], o[s])
        assert errs["times"] == 0.0
        assert compare.worst(errs) <= tol, errs
        pe = compare.pointwise_errors(e[s], o[s])
        assert compare.worst(pe) <= 100 * tol, pe


def test_swap_rate_annuity_basis_changes_the_fit(emu):  # noqa: F811
    """The two bases give different continuation values (the engine honours the choice)."""
    spec = W.c3(num_paths=400, dates=30, num_swaptions=12)
    try:
        _set_basis(emu, abi.CCR_BASIS_STATE)
        a = run_emu(emu, spec)[0].get("expected_exposure")
        _set_basis(emu, abi.CCR_BASIS_SWAP_RATE_ANNUITY)
        b = run_emu(emu, spec)[0].get("expected_exposure")
    finally:
        _set_basis(emu, -1)
    rel = np.abs(a - b).max() / np.abs(a).max()
    assert 1e-6 < rel < 0.2, rel
