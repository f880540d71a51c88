"""Seeded synthetic workloads: the five benchmark portfolios and small parity cases.

Trades are sampled with SplitMix64 (seed per workload) so every process, here or on the
GPU host, builds bit-identical inputs. Sizes (paths, dates) can be scaled down for parity
tests; `full=True` gives the benchmark shapes.

  c1  Plain FP64, one uncollateralised netting set of 10 vanilla IRS (semiannual fixed vs
      quarterly float), HW1F a=0.03 sigma=0.01 on a flat 3% curve, 5,000 paths x 120 monthly
      dates, PFE 97.5%.
  c2  3 currencies on Hull-White plus 2 lognormal FX rates with a random SPD correlation
      (|rho| <= 0.6); 1,000 trades (700 IRS, 300 FX forwards) over 20 netting sets under CSA
      (threshold 0-1e7, MTA 5e5, 10-day MPoR); 50,000 paths x 260 weekly dates.
  c3  200 Bermudan payer/receiver swaptions on HW1F (AMC), 100,000 paths x 120 dates.
  c4  HW + 1 FX + CIR++ counterparty/own intensities with 0.3 wrong-way correlation.
  c5  500 netting sets, Zipf(1.2) trade counts, half under CSA.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from .abi import (CCR_COLLATERAL_ASSET, CCR_COLLATERAL_BOND, CCR_OPTION_CALL, CCR_OPTION_PUT,
                  CCR_TRADE_ASSET_FORWARD, CCR_TRADE_BERMUDAN_SWAPTION, CCR_TRADE_EUROPEAN_OPTION, CCR_TRADE_SWAP,
                  Asset, Cir, Collateral, CreditCurve, Curve, HullWhite, Model, NettingSet, Spec, Trade, swap_schedule, uniform_grid)

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): state += 0x9e3779b97f4a7c15, then two xor-shift-multiply rounds."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def bernoulli(self) -> bool:
        return (self.next() >> 63) == 1


def two_leg_swap(rng: SplitMix64, notional_lo=1e6, notional_hi=1e8, mat_lo=1.0, mat_hi=10.0, currency=0) -> Trade:
    notional = rng.uniform(notional_lo, notional_hi)
    pay = rng.bernoulli()
    maturity = rng.uniform(mat_lo, mat_hi)
    k = rng.uniform(0.02, 0.04)
    return Trade(CCR_TRADE_SWAP, notional, k, pay, schedule=swap_schedule(0.0, maturity, 2),
                 float_schedule=swap_schedule(0.0, maturity, 4), currency=currency)


def random_correlation(rng: SplitMix64, n: int, max_abs: float = 0.6) -> np.ndarray:
    """Seeded SPD correlation with |rho| <= max_abs: normalised Gram of random vectors, rejected until it fits."""
    while True:
        v = np.array([[rng.uniform(-1.0, 1.0) for _ in range(n + 2)] for _ in range(n)])
        g = v @ v.T
        d = np.sqrt(np.diag(g))
        c = g / np.outer(d, d)
        np.fill_diagonal(c, 1.0)
        off = np.abs(c[~np.eye(n, dtype=bool)])
        if off.size == 0 or off.max() <= max_abs:
            if np.linalg.eigvalsh(c).min() > 1e-6:
                return c


def c1(num_paths: int = 5000, dates: int = 120, trades: int = 10, seed: int = 42) -> Spec:
    rng = SplitMix64(1)
    model = Model(HullWhite(Curve.flat(0.03), 0.03, 0.01))
    ts = [two_leg_swap(rng) for _ in range(trades)]
    return Spec(model, [NettingSet(ts)], uniform_grid(10.0, dates), num_paths, seed, True, 0.975, "c1")


def c2_model(rng: SplitMix64) -> Model:
    rates = [HullWhite(Curve.flat(rng.uniform(0.01, 0.04)), rng.uniform(0.01, 0.1), rng.uniform(0.005, 0.015))
             for _ in range(3)]
    spots = [1.1, 0.9]
    assets = [Asset(spots[i], rng.uniform(0.08, 0.15), rates[1 + i].curve, foreign_rate=i) for i in range(2)]
    corr = random_correlation(rng, 5, 0.6)
    return Model(rates[0], rates[1:], assets, [], corr)


def c2(num_paths: int = 50000, weeks: int = 260, num_sets: int = 20, num_trades: int = 1000,
       seed: int = 42, horizon: float = 5.0) -> Spec:
    rng = SplitMix64(2)
    model = c2_model(rng)
    n_irs = int(round(num_trades * 0.7))
    trades: List[Trade] = []
    for i in range(num_trades):
        if i < n_irs:
            ccy = rng.integer(0, 2)
            trades.append(two_leg_swap(rng, mat_lo=0.5, mat_hi=horizon, currency=ccy))
        else:
            a = rng.integer(0, 1)
            n = rng.uniform(1e6, 1e8) * (1.0 if rng.bernoulli() else -1.0)
            mat = rng.uniform(0.25, horizon)
            fwd = model.assets[a].spot * math.exp((model.rates.curve.rates[0] - model.assets[a].carry.rates[0]) * mat)
            trades.append(Trade(CCR_TRADE_ASSET_FORWARD, n, fwd * rng.uniform(0.9, 1.1), asset=a, maturity=mat))
    # Deal IRS and forwards round-robin so every set mixes both kinds.
    order = [i for pair in zip(range(n_irs), range(n_irs, num_trades)) for i in pair]
    order += [i for i in range(num_trades) if i not in set(order)]
    sets = [NettingSet([]) for _ in range(num_sets)]
    for k, i in enumerate(order):
        sets[k % num_sets].trades.append(trades[i])
    for ns in sets:
        th = rng.uniform(0.0, 1e7)
        ns.collateral = Collateral(threshold_counterparty=th, threshold_own=th, mta=5e5, mpr=10.0 / 250.0)
    return Spec(model, sets, uniform_grid(horizon, weeks), num_paths, seed, True, 0.975, "c2")


def bermudan(rng: SplitMix64, notional_hi=5e7) -> Trade:
    notional = rng.uniform(1e6, notional_hi)
    pay = rng.bernoulli()
    expiry = float(rng.integer(1, 10))
    tenor = float(rng.integer(5, 10))
    sched = swap_schedule(expiry, tenor, 1)
    return Trade(CCR_TRADE_BERMUDAN_SWAPTION, notional, rng.uniform(0.02, 0.04), pay, schedule=sched,
                 exercise=sched[:-1].copy(), position=1.0, degree=3)


def c3(num_paths: int = 100000, dates: int = 120, num_swaptions: int = 200, seed: int = 42) -> Spec:
    rng = SplitMix64(3)
    model = Model(HullWhite(Curve.flat(0.03), 0.03, 0.01))
    ts = [bermudan(rng) for _ in range(num_swaptions)]
    horizon = max(t.schedule[-1] for t in ts)
    return Spec(model, [NettingSet(ts)], uniform_grid(horizon, dates), num_paths, seed, True, 0.975, "c3")


def c4_model(rho_wrong_way: float = 0.3) -> Model:
    dom = HullWhite(Curve.flat(0.03), 0.03, 0.01)
    fx = Asset(1.1, 0.12, Curve.flat(0.02))
    hazard = CreditCurve.flat(0.02)
    cpty = Cir(hazard, 0.5, 0.02, 0.1, 0.02, 4)
    own = Cir(CreditCurve.flat(0.01), 0.5, 0.01, 0.05, 0.01, 4)
    corr = np.eye(4)
    corr[1, 2] = corr[2, 1] = rho_wrong_way  # FX driver vs counterparty intensity
    return Model(dom, [], [fx], [cpty, own], corr)


def c4(num_paths: int = 100000, dates: int = 120, seed: int = 42, with_bermudans: bool = True) -> Spec:
    rng = SplitMix64(4)
    model = c4_model()
    ts: List[Trade] = [two_leg_swap(rng) for _ in range(20)]
    for _ in range(10):
        n = rng.uniform(1e6, 1e8)
        mat = rng.uniform(1.0, 10.0)
        ts.append(Trade(CCR_TRADE_ASSET_FORWARD, n, 1.1 * rng.uniform(0.9, 1.1), asset=0, maturity=mat))
    if with_bermudans:
        ts += [bermudan(rng) for _ in range(10)]
    horizon = max(max(t.schedule[-1] if t.schedule is not None else t.maturity for t in ts), 10.0)
    ns = NettingSet(ts, counterparty_credit=0, own_credit=1)
    return Spec(model, [ns], uniform_grid(horizon, dates), num_paths, seed, True, 0.975, "c4")


def c5(num_paths: int = 20000, dates: int = 120, num_sets: int = 500, largest: int = 5000, seed: int = 42,
       bermudan_share: float = 0.05) -> Spec:
    """Batched uneven portfolio: num_sets netting sets with round(largest / rank^1.2) trades each
    (Zipf s = 1.2, clipped to [1, largest]; ~20,000 trades at 500 sets), a mix of two-leg IRS in
    3 currencies, FX forwards and Bermudan swaptions, every other set under CSA; 3-currency
    Hull-White + 2 FX + one CIR++ intensity shared by the counterparties."""
    rng = SplitMix64(5)
    base = c2_model(rng)
    corr = np.eye(6)
    corr[:5, :5] = base.correlation
    cir = Cir(CreditCurve.flat(0.02), 0.5, 0.02, 0.1, 0.02, 4)
    model = Model(base.rates, base.foreign, base.assets, [cir], corr)
    sets = []
    for i in range(1, num_sets + 1):
        n = max(1, min(largest, int(round(largest / i ** 1.2))))
        ts: List[Trade] = []
        for _ in range(n):
            u = rng.uniform()
            if u < bermudan_share:
                ts.append(bermudan(rng))
            elif u < bermudan_share + 0.3:
                a = rng.integer(0, 1)
                notional = rng.uniform(1e6, 1e8) * (1.0 if rng.bernoulli() else -1.0)
                mat = 0.25 * rng.integer(1, 40)
                fwd = model.assets[a].spot * math.exp((model.rates.curve.rates[0] - model.assets[a].carry.rates[0]) * mat)
                ts.append(Trade(CCR_TRADE_ASSET_FORWARD, notional, fwd * rng.uniform(0.9, 1.1), asset=a, maturity=mat))
            else:
                mat = 0.25 * rng.integer(2, 40)
                ts.append(Trade(CCR_TRADE_SWAP, rng.uniform(1e6, 1e8), rng.uniform(0.02, 0.04), rng.bernoulli(),
                                schedule=swap_schedule(0.0, mat, 2), float_schedule=swap_schedule(0.0, mat, 4),
                                currency=rng.integer(0, 2)))
        ns = NettingSet(ts, counterparty_credit=0)
        if i % 2 == 0:
            th = rng.uniform(0.0, 1e7)
            ns.collateral = Collateral(threshold_counterparty=th, threshold_own=th, mta=5e5, mpr=10.0 / 250.0)
        sets.append(ns)
    horizon = 20.0
    return Spec(model, sets, uniform_grid(horizon, dates), num_paths, seed, True, 0.975, "c5")


def small_reference_case(num_paths: int = 64, dates: int = 12, csa: bool = False, fx: bool = False,
                         credits: bool = False, two_leg: bool = False, seed: int = 7, bermudans: int = 0,
                         cash_settled: bool = False, options: int = 0, security: str = "cash") -> Spec:
    """Small reference-representable netting set for parity against a build of the reference.
    security (with csa): the collateral posted -- "cash", "bond" (a 6y zero bond at a 2%
    haircut) or "asset" (the FX asset at 8%; needs fx)."""
    rng = SplitMix64(11)
    curve = Curve([0.5, 2.0, 5.0, 10.0], [0.020, 0.025, 0.030, 0.033])
    model = Model(HullWhite(curve, 0.05, 0.01))
    nf = 1
    if fx:
        model.assets.append(Asset(1.1, 0.12, Curve.flat(0.01)))
        nf += 1
    if credits:
        model.credits.append(Cir(CreditCurve([1.0, 3.0, 5.0], [0.01, 0.015, 0.02]), 0.5, 0.012, 0.08, 0.008, 4))
        model.credits.append(Cir(CreditCurve.flat(0.01), 0.4, 0.01, 0.05, 0.01, 2))
        nf += 2
    if nf > 1:
        model.correlation = random_correlation(SplitMix64(5), nf, 0.5)
    ts = []
    for i in range(4):
        if two_leg:
            ts.append(two_leg_swap(rng, mat_lo=1.0, mat_hi=5.0))
        else:
            mat = rng.uniform(1.0, 5.0)
            ts.append(Trade(CCR_TRADE_SWAP, rng.uniform(1e5, 1e6), rng.uniform(0.02, 0.035), rng.bernoulli(),
                            schedule=swap_schedule(0.0, mat, 2)))
    if fx:
        ts.append(Trade(CCR_TRADE_ASSET_FORWARD, 5e5, 1.05, asset=0, maturity=3.0))
        ts.append(Trade(CCR_TRADE_ASSET_FORWARD, -3e5, 1.12, asset=0, maturity=1.5))
    for i in range(options if fx else 0):
        # FX calls and puts, long and short, 0.5-3y expiries.
        ts.append(Trade(CCR_TRADE_EUROPEAN_OPTION, rng.uniform(1e5, 1e6) * (-1.0 if i % 3 == 2 else 1.0),
                        rng.uniform(0.95, 1.25), asset=0, maturity=rng.uniform(0.5, 3.0),
                        option_type=CCR_OPTION_PUT if i % 2 else CCR_OPTION_CALL))
    for i in range(bermudans):
        # 1-2y into 2-3y, semiannual exercise on the underlying's period starts.
        expiry = 1.0 + 0.5 * rng.integer(0, 2)
        sched = swap_schedule(expiry, 2.0 + 0.5 * rng.integer(0, 2), 2)
        ts.append(Trade(CCR_TRADE_BERMUDAN_SWAPTION, rng.uniform(1e5, 1e6), rng.uniform(0.02, 0.035), rng.bernoulli(),
                        schedule=sched, exercise=sched[:-1].copy(), position=-1.0 if i % 2 else 1.0,
                        degree=1 + i % 3, cash_settled=cash_settled))
    ns = NettingSet(ts)
    if csa:
        ns.collateral = Collateral(threshold_counterparty=2e4, threshold_own=5e4, mta=1e4, mpr=10.0 / 250.0)
        if security == "bond":
            ns.collateral.type, ns.collateral.haircut, ns.collateral.bond_maturity = CCR_COLLATERAL_BOND, 0.02, 6.0
        elif security == "asset":
            ns.collateral.type, ns.collateral.haircut, ns.collateral.asset_index = CCR_COLLATERAL_ASSET, 0.08, 0
        elif security != "cash":
            raise ValueError(f"unknown collateral security {security!r}")
    if credits:
        ns.counterparty_credit, ns.own_credit = 0, 1
    return Spec(model, [ns], uniform_grid(5.0, dates), num_paths, seed, True, 0.95, "small")


def book_case(num_paths: int = 256, dates: int = 10, seed: int = 7, secured: bool = False) -> Spec:
    """A small portfolio book: three netting sets on different date sets (a CSA set of swaps, an
    uncollateralised set of swaps and FX forwards, an uncollateralised set of FX options), the
    first two facing counterparty credit 0 and the third credit 1. secured: the second and third
    sets are under CSAs too, collateralised with a zero bond and with the FX asset."""
    rng = SplitMix64(23)
    curve = Curve([0.5, 2.0, 5.0, 10.0], [0.020, 0.025, 0.030, 0.033])
    model = Model(HullWhite(curve, 0.05, 0.01))
    model.assets.append(Asset(1.1, 0.12, Curve.flat(0.01)))
    model.credits.append(Cir(CreditCurve([1.0, 3.0, 5.0], [0.01, 0.015, 0.02]), 0.5, 0.012, 0.08, 0.008, 4))
    model.credits.append(Cir(CreditCurve.flat(0.01), 0.4, 0.01, 0.05, 0.01, 2))
    model.correlation = random_correlation(SplitMix64(5), 4, 0.5)
    swaps = [Trade(CCR_TRADE_SWAP, rng.uniform(1e5, 1e6), rng.uniform(0.02, 0.035), rng.bernoulli(),
                   schedule=swap_schedule(0.0, rng.uniform(1.0, 5.0), 2)) for _ in range(3)]
    a = NettingSet(swaps)
    a.collateral = Collateral(threshold_counterparty=2e4, threshold_own=5e4, mta=1e4, mpr=10.0 / 250.0)
    a.counterparty_credit, a.own_credit = 0, 1
    b = NettingSet([Trade(CCR_TRADE_SWAP, rng.uniform(1e5, 1e6), rng.uniform(0.02, 0.035), rng.bernoulli(),
                          schedule=swap_schedule(0.0, rng.uniform(1.0, 4.0), 4)),
                    Trade(CCR_TRADE_ASSET_FORWARD, 5e5, 1.05, asset=0, maturity=2.7),
                    Trade(CCR_TRADE_ASSET_FORWARD, -3e5, 1.12, asset=0, maturity=1.3)])
    b.counterparty_credit, b.own_credit = 0, 1
    c = NettingSet([Trade(CCR_TRADE_EUROPEAN_OPTION, rng.uniform(1e5, 1e6) * (-1.0 if i == 2 else 1.0),
                          rng.uniform(0.95, 1.25), asset=0, maturity=rng.uniform(0.5, 3.0),
                          option_type=CCR_OPTION_PUT if i % 2 else CCR_OPTION_CALL) for i in range(3)])
    c.counterparty_credit, c.own_credit = 1, 0
    if secured:
        b.collateral = Collateral(threshold_counterparty=1e4, threshold_own=3e4, mta=5e3, mpr=10.0 / 250.0,
                                  type=CCR_COLLATERAL_BOND, haircut=0.03, bond_maturity=7.0)
        c.collateral = Collateral(threshold_counterparty=1e4, threshold_own=2e4, mta=5e3, mpr=10.0 / 250.0,
                                  type=CCR_COLLATERAL_ASSET, haircut=0.1, asset_index=0)
    return Spec(model, [a, b, c], uniform_grid(5.0, dates), num_paths, seed, True, 0.95, "book")
