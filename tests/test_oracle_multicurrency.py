"""The oracle's multi-currency extension, pinned without a reference oracle (CPU only).

The reference simulates one rate factor; foreign rates there are deterministic carry curves
(include/ccr/models/lognormal_asset.hpp:10-16, src/models/lognormal_asset.cpp:17-22). The oracle
adds foreign Hull-White rates with the quanto drift -rho(x_f, S) sigma_f sigma_S under the
domestic measure, FX carried by the simulated foreign short rate, and foreign swaps converted at
the simulated spot. Those parts are pinned here by:

1. the sigma_f = 0 reduction: a foreign rate with no volatility is the reference's
   deterministic-carry model; on the same FX draws (replayed under both draw layouts) spots,
   deflators and whole exposure profiles are bit-identical;
2. Monte Carlo martingales of the domestic risk-neutral measure, within 4 standard errors (model
   after tests/test_models.cpp:92-108 of the reference): E[D S_a(t)] = S_a(0) P_f(0,t) and
   E[D S_a(t) P_f(t,T|x_f)] = S_a(0) P_f(0,T) on the scenarios; on the exposure profile,
   E[D V(t)] of a foreign swap equals spot x its foreign value at 0 of the periods still unpaid
   at t (exactly spot x npv at t = 0), and E[D V(t)] of an FX forward equals
   N (S_a(0) P_f(0,T) - K P_d(0,T));
3. a mutation check: the oracle built with the quanto drift's sign flipped (oracle/Makefile
   `mutants`) fails the martingale checks by a wide margin.
"""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_lib as OL
from zlatko_amd import abi
from zlatko_amd.abi import (CCR_TRADE_ASSET_FORWARD, CCR_TRADE_SWAP, Asset, Collateral, Curve, HullWhite, Model,
                            NettingSet, Spec, Trade, swap_schedule, uniform_grid)

Z_BOUND = 4.0


def quanto_model() -> Model:
    """Domestic + two foreign Hull-White rates, two FX spots converting them, strong quanto
    correlations (|rho(x_f, S_f)| = 0.5) so a drift error is visible within a few standard errors."""
    dom = HullWhite(Curve([1.0, 5.0, 10.0], [0.02, 0.025, 0.03]), 0.05, 0.01)
    f0 = HullWhite(Curve.flat(0.035), 0.03, 0.015)
    f1 = HullWhite(Curve([2.0, 10.0], [0.01, 0.015]), 0.08, 0.012)
    assets = [Asset(1.1, 0.15, f0.curve, foreign_rate=0), Asset(0.9, 0.12, f1.curve, foreign_rate=1)]
    # Factor order [domestic, foreign 0, foreign 1, spot 0, spot 1].
    c = np.eye(5)
    for (i, j), r in {(0, 1): 0.3, (0, 2): 0.2, (0, 3): 0.2, (0, 4): -0.1, (1, 2): 0.25, (1, 3): 0.5, (1, 4): 0.1,
                      (2, 3): 0.0, (2, 4): -0.5, (3, 4): 0.3}.items():
        c[i, j] = c[j, i] = r
    assert np.linalg.eigvalsh(c).min() > 0.05
    return Model(dom, [f0, f1], assets, [], c)


def _pair_mean_se(m: np.ndarray):
    """Mean and standard error over antithetic pairs (paths 2q, 2q + 1 share their draws)."""
    pairs = 0.5 * (m[0::2] + m[1::2])
    return pairs.mean(), pairs.std(ddof=1) / np.sqrt(len(pairs))


def scenario_martingale_z(lib, num_paths: int = 20000) -> float:
    """Worst |z| over dates of E[D], E[D P_d(t,12)], E[D S_a], E[D S_a P_f(t,12)] against their
    time-0 values on the oracle's scenarios."""
    m = quanto_model()
    grid = uniform_grid(10.0, 20)
    spec = Spec(m, [NettingSet([Trade(CCR_TRADE_ASSET_FORWARD, 1.0, maturity=1.0)])], grid, num_paths, seed=5)
    sc = OL.scenarios(spec, grid, lib=lib)
    T = 12.0
    worst = 0.0
    for j in range(1, len(grid)):
        t = grid[j]
        D = sc["D"][:, j]
        checks = [(D, OL.discount(m.rates.curve, t))]
        a, b = OL.zero_bond_factors(m.rates, t, T)
        checks.append((D * a * np.exp(-b * sc["x"][:, j]), OL.discount(m.rates.curve, T)))
        for k, asset in enumerate(m.assets):
            f = m.foreign[asset.foreign_rate]
            S = sc["S"][k][:, j]
            checks.append((D * S, asset.spot * OL.discount(f.curve, t)))
            a, b = OL.zero_bond_factors(f, t, T)
            pf = a * np.exp(-b * sc["xf"][asset.foreign_rate][:, j])
            checks.append((D * S * pf, asset.spot * OL.discount(f.curve, T)))
        for v, want in checks:
            mean, se = _pair_mean_se(v)
            worst = max(worst, abs(mean - want) / se)
    return worst


def _foreign_swap(pay: bool) -> Trade:
    return Trade(CCR_TRADE_SWAP, 1e7, 0.035, pay, schedule=swap_schedule(0.0, 8.0, 2),
                 float_schedule=swap_schedule(0.0, 8.0, 4), currency=1)


def _valuation_spec(num_paths: int) -> Spec:
    sets = [NettingSet([_foreign_swap(True)]), NettingSet([_foreign_swap(False)]),
            NettingSet([Trade(CCR_TRADE_ASSET_FORWARD, 1e6, 0.0, asset=0, maturity=8.0)]),
            NettingSet([Trade(CCR_TRADE_ASSET_FORWARD, -2e6, 0.95, asset=1, maturity=6.0)])]
    return Spec(quanto_model(), sets, uniform_grid(10.0, 20), num_paths, seed=9, pfe_quantile=0.95)


def _unpaid_foreign_value(tr: Trade, f: HullWhite, t: float) -> float:
    """Foreign value at time 0 of the periods of a two-leg swap that are unpaid at t
    (te > t + 1e-10): floating N s (P(0,ts) - P(0,te)), fixed -N s K tau P(0,te)."""
    s = 1.0 if tr.pay_fixed else -1.0
    v = 0.0
    fl, fx = tr.float_schedule, tr.schedule
    for ts, te in zip(fl[:-1], fl[1:]):
        if te > t + 1e-10:
            v += s * tr.notional * (OL.discount(f.curve, ts) - OL.discount(f.curve, te))
    for ts, te in zip(fx[:-1], fx[1:]):
        if te > t + 1e-10:
            v -= s * tr.notional * tr.fixed_rate * (te - ts) * OL.discount(f.curve, te)
    return v


def valuation_martingale_z(lib, num_paths: int = 20000):
    """(worst |z| of E[D V(t)] against its martingale value over dates and trades, worst relative
    error of V(0) against spot x npv) from the oracle's exposure profiles and exposure rows."""
    spec = _valuation_spec(num_paths)
    outs = OL.run_oracle_with(spec, lib, threads=8, rows=True)
    m = spec.model
    sim = OL.simulation_grid(spec, 0)
    D = OL.scenarios(spec, sim, lib=lib)["D"]
    P = spec.num_paths
    worst, t0err = 0.0, 0.0
    swap = spec.sets[0].trades[0]
    f = m.foreign[0]
    spot = m.assets[0].spot  # asset 0 converts foreign rate 0 (the swap's currency 1)
    rows = [o.exposure_rows(P) for o in outs]
    times = outs[0].get("times")
    for r, t in enumerate(times):
        j = int(np.argmin(np.abs(sim - t)))
        assert abs(sim[j] - t) <= 1e-12
        v = rows[0][r] - rows[1][r]  # the receiver swap is the exact negative of the payer swap
        want = spot * _unpaid_foreign_value(swap, f, t)
        if r == 0:
            t0err = abs(v[0] - want) / (spot * swap.notional)
            continue
        if t >= 8.0 - 1e-10:
            continue
        mean, se = _pair_mean_se(D[:, j] * v)
        dev = outs[0].get("discounted_expected_exposure")[r] + outs[0].get("discounted_expected_negative_exposure")[r]
        assert abs(dev - mean) <= 1e-9 * abs(spot * swap.notional)  # the profile's E[D V] is that same mean
        worst = max(worst, abs(mean - want) / se)
    for k, s in ((2, 0), (3, 1)):
        fwd = spec.sets[k].trades[0]
        asset = m.assets[fwd.asset]
        fr = m.foreign[asset.foreign_rate]
        o = outs[k]
        tk = o.get("times")
        simk = OL.simulation_grid(spec, k)
        Dk = OL.scenarios(spec, simk, lib=lib)["D"]
        rk = o.exposure_rows(P)
        want = fwd.notional * (asset.spot * OL.discount(fr.curve, fwd.maturity)
                               - fwd.fixed_rate * OL.discount(m.rates.curve, fwd.maturity))
        for r, t in enumerate(tk):
            if r == 0 or t >= fwd.maturity - 1e-10:
                continue
            j = int(np.argmin(np.abs(simk - t)))
            dv = o.get("discounted_expected_exposure")[r] + o.get("discounted_expected_negative_exposure")[r]
            if fwd.fixed_rate == 0.0:  # V > 0 on every path: the row is V itself
                mean, se = _pair_mean_se(Dk[:, j] * rk[r])
                assert abs(dv - mean) <= 1e-9 * abs(want)
            else:  # standard error bounded by that of D |V| <= D (N S P_f + |N| K)
                _, se = _pair_mean_se(Dk[:, j] * rk[r])
                se = max(se, 1e-3 * abs(want) / np.sqrt(P))
            worst = max(worst, abs(dv - want) / se)
    return worst, t0err


def test_scenario_martingales_of_the_three_currency_model():
    z = scenario_martingale_z(OL.oracle())
    print(f"scenario martingales: worst |z| {z:.2f}", flush=True)
    assert z < Z_BOUND


def test_foreign_swap_and_fx_forward_values_are_martingales():
    z, t0 = valuation_martingale_z(OL.oracle())
    print(f"valuation martingales: worst |z| {z:.2f}, t = 0 foreign swap vs spot x npv {t0:.2e}", flush=True)
    assert z < Z_BOUND and t0 < 1e-12


def test_quanto_sign_mutant_is_rejected():
    """The checks above are not vacuous: flipping the sign of the quanto drift in the oracle (the
    mutant build) moves the foreign-measure martingales far outside their bounds."""
    lib = OL.mutant_quanto()
    zs = scenario_martingale_z(lib)
    zv, _ = valuation_martingale_z(lib)
    print(f"quanto-sign mutant: scenario |z| {zs:.1f}, valuation |z| {zv:.1f}", flush=True)
    assert zs > 3 * Z_BOUND and zv > 3 * Z_BOUND


def _carry_pair():
    """The reference model (domestic HW + one FX spot on a deterministic carry curve) and the same
    market written with a zero-volatility foreign Hull-White rate on that curve."""
    carry = Curve([1.0, 4.0, 8.0], [0.01, 0.015, 0.02])
    dom = HullWhite(Curve([1.0, 5.0, 10.0], [0.02, 0.025, 0.03]), 0.05, 0.01)
    rho = 0.35
    a = Model(dom, [], [Asset(1.1, 0.12, carry)], [], np.array([[1.0, rho], [rho, 1.0]]))
    cb = np.eye(3)
    cb[0, 2] = cb[2, 0] = rho  # [domestic, foreign (independent), spot]
    b = Model(dom, [HullWhite(carry, 0.05, 0.0)], [Asset(1.1, 0.12, carry, foreign_rate=0)], [], cb)
    trades = [Trade(CCR_TRADE_ASSET_FORWARD, 5e5, 1.05, asset=0, maturity=3.0),
              Trade(CCR_TRADE_ASSET_FORWARD, -3e5, 1.12, asset=0, maturity=1.5),
              Trade(CCR_TRADE_SWAP, 4e5, 0.025, True, schedule=swap_schedule(0.0, 4.0, 2),
                    float_schedule=swap_schedule(0.0, 4.0, 4))]
    ns = NettingSet(trades, collateral=Collateral(threshold_counterparty=2e4, threshold_own=5e4, mta=1e4))
    return a, b, ns


def test_zero_vol_foreign_rate_is_the_carry_curve_path_by_path():
    """sigma_f = 0: the foreign state stays 0 with no quanto drift and its integrated short rate
    is the carry integral ln(P(t0)/P(t1)) exactly, so on the same FX and domestic draws (the
    foreign draws are arbitrary) spots, deflators and every exposure-profile field and exposure
    row are bit-identical to the deterministic-carry model's."""
    ma, mb, ns = _carry_pair()
    P = 400
    grid = uniform_grid(5.0, 20)
    sa = Spec(ma, [ns], grid, P, seed=3)
    sb = Spec(mb, [ns], grid, P, seed=3)
    sim = OL.simulation_grid(sa, 0)
    assert np.array_equal(sim, OL.simulation_grid(sb, 0))
    steps = len(sim) - 1
    rng = np.random.default_rng(17)
    za = rng.standard_normal((P // 2, steps, 3))  # [domestic, spot | domestic integral]
    zb = rng.standard_normal((P // 2, steps, 5))  # [domestic, foreign, spot | domestic integral, foreign integral]
    zb[..., 0], zb[..., 2], zb[..., 3] = za[..., 0], za[..., 1], za[..., 2]
    sca, scb = OL.scenarios(sa, sim, za), OL.scenarios(sb, sim, zb)
    assert np.array_equal(sca["S"][0], scb["S"][0])
    assert np.array_equal(sca["D"], scb["D"]) and np.array_equal(sca["x"], scb["x"])
    assert not np.any(scb["xf"][0])
    oa = OL.run_oracle_replay(sa, za, rows=True)[0]
    ob = OL.run_oracle_replay(sb, zb, rows=True)[0]
    for fld in abi.PROFILE_FIELDS:
        assert np.array_equal(oa.get(fld), ob.get(fld)), fld
    assert np.array_equal(oa.exposure_rows(P), ob.exposure_rows(P))
    # The replayed normals drive the engine: other draws give another profile.
    oc = OL.run_oracle_replay(sa, -za, rows=False)[0]
    assert not np.array_equal(oa.get("expected_exposure"), oc.get("expected_exposure"))


def test_replayed_stream_reproduces_the_seeded_run():
    """Replaying the seeded stream's own normals reproduces the seeded run bit for bit (the
    replay consumes draws in the generator's order)."""
    m = quanto_model()
    ns = NettingSet([_foreign_swap(True), Trade(CCR_TRADE_ASSET_FORWARD, 1e6, 1.0, asset=1, maturity=4.0)])
    spec = Spec(m, [ns], uniform_grid(5.0, 10), 64, seed=21)
    steps = len(OL.simulation_grid(spec, 0)) - 1
    z = OL.normals(21, 32 * steps * m.num_normals())
    a = OL.run_oracle(spec, rows=True)[0]
    b = OL.run_oracle_replay(spec, z, rows=True)[0]
    for fld in abi.PROFILE_FIELDS:
        assert np.array_equal(a.get(fld), b.get(fld)), fld
