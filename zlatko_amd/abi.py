"""ctypes mirror of include/ccr_hip.h and Python builders for its input structs.

The structs are plain C (no torch types); a `Spec` keeps every numpy buffer alive while
the C structs point into it. The same structs feed the product library (libccr_hip.so)
and, in tests, the FP64 oracle.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

DP = C.POINTER(C.c_double)

CCR_OK = 0
CCR_ERR_INVALID_ARGUMENT = 1
CCR_ERR_OUT_OF_RANGE = 2
CCR_ERR_UNSUPPORTED = 3
CCR_ERR_OUT_OF_MEMORY = 4
CCR_ERR_DEVICE = 5
CCR_ERR_RUNTIME = 6

CCR_TRADE_SWAP = 0
CCR_TRADE_ASSET_FORWARD = 1
CCR_TRADE_BERMUDAN_SWAPTION = 2
CCR_TRADE_EUROPEAN_OPTION = 3
CCR_OPTION_CALL = 0
CCR_OPTION_PUT = 1

CCR_COLLATERAL_CASH = 0
CCR_COLLATERAL_BOND = 1
CCR_COLLATERAL_ASSET = 2

CCR_FP64 = 0
CCR_FP32 = 1

CCR_BASIS_STATE = 0
CCR_BASIS_SWAP_RATE_ANNUITY = 1

CCR_FAULT_NONE = 0
CCR_FAULT_ALLOC = 1
CCR_FAULT_LAUNCH = 2
CCR_FAULT_PLAN = 3


class YieldCurveC(C.Structure):
    _fields_ = [("times", DP), ("zero_rates", DP), ("size", C.c_int32)]


class CreditCurveC(C.Structure):
    _fields_ = [("pillars", DP), ("hazards", DP), ("size", C.c_int32)]


class XvaParametersC(C.Structure):
    _fields_ = [("funding_spread_borrow", C.c_double), ("funding_spread_lend", C.c_double),
                ("margin_funding_spread", C.c_double), ("hurdle_rate", C.c_double), ("capital_ratio", C.c_double),
                ("risk_weight", C.c_double), ("alpha", C.c_double)]


class XvaResultC(C.Structure):
    _fields_ = [("cva", C.c_double), ("dva", C.c_double), ("fca", C.c_double), ("fba", C.c_double),
                ("mva", C.c_double), ("kva", C.c_double)]


TRADE_STATS = ("marginal_discounted", "marginal", "incremental_discounted", "standalone_discounted", "standalone")


class TradeStatisticsC(C.Structure):
    _fields_ = [("capacity", C.c_int32), ("num_dates", C.c_int32)] + [(n, DP) for n in TRADE_STATS]


class HullWhiteC(C.Structure):
    _fields_ = [("curve", YieldCurveC), ("mean_reversion", C.c_double), ("volatility", C.c_double)]


class AssetC(C.Structure):
    _fields_ = [("spot", C.c_double), ("volatility", C.c_double), ("carry", YieldCurveC),
                ("foreign_rate", C.c_int32)]


class CirC(C.Structure):
    _fields_ = [("market", CreditCurveC), ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double),
                ("y0", C.c_double), ("substeps", C.c_int32)]


class GeneratorC(C.Structure):
    _fields_ = [("rates", HullWhiteC), ("num_foreign_rates", C.c_int32), ("foreign_rates", C.POINTER(HullWhiteC)),
                ("num_assets", C.c_int32), ("assets", C.POINTER(AssetC)), ("num_credits", C.c_int32),
                ("credits", C.POINTER(CirC)), ("correlation", DP)]


class SimConfigC(C.Structure):
    _fields_ = [("num_paths", C.c_int64), ("seed", C.c_uint64), ("antithetic", C.c_int32)]


class TradeC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pay_fixed", C.c_int32), ("currency", C.c_int32), ("asset", C.c_int32),
                ("notional", C.c_double), ("fixed_rate", C.c_double), ("maturity", C.c_double),
                ("schedule", DP), ("schedule_size", C.c_int32), ("float_schedule", DP),
                ("float_schedule_size", C.c_int32), ("exercise_times", DP), ("num_exercises", C.c_int32),
                ("cash_settled", C.c_int32), ("position", C.c_double), ("regression_degree", C.c_int32),
                ("option_type", C.c_int32)]


class CollateralC(C.Structure):
    _fields_ = [("threshold_counterparty", C.c_double), ("threshold_own", C.c_double),
                ("minimum_transfer_amount", C.c_double), ("independent_amount", C.c_double),
                ("margin_period_of_risk", C.c_double), ("type", C.c_int32), ("haircut", C.c_double),
                ("bond_maturity", C.c_double), ("asset_index", C.c_int32)]


class NettingSetC(C.Structure):
    _fields_ = [("num_trades", C.c_int32), ("trades", C.POINTER(TradeC)), ("has_collateral", C.c_int32),
                ("collateral", CollateralC), ("counterparty_credit", C.c_int32), ("own_credit", C.c_int32)]


class ExposureOutC(C.Structure):
    _fields_ = [("capacity", C.c_int32), ("num_dates", C.c_int32)] + [
        (n, DP) for n in ("times", "expected_value", "expected_exposure", "expected_negative_exposure",
                          "potential_future_exposure", "discounted_expected_exposure",
                          "discounted_expected_negative_exposure", "mean_deflator", "mean_survival",
                          "cva_increments", "dva_increments", "exposure_rows")]


class HipOptionsC(C.Structure):
    _fields_ = [("device", C.c_int32), ("precision", C.c_int32), ("block_size", C.c_int32)]


class HipTimingC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("path_ms", "reduce_ms", "pfe_ms", "total_ms", "upload_ms", "draws_ms",
                                          "amc_ms")]


class HipPlanInfoC(C.Structure):
    _fields_ = [("num_sets", C.c_int32), ("num_paths", C.c_int64), ("max_sim_dates", C.c_int32),
                ("max_reporting_dates", C.c_int32), ("num_normals", C.c_int32), ("total_terms", C.c_int64),
                ("stream_normals", C.c_int64), ("revaluations", C.c_int64), ("term_evaluations", C.c_int64),
                ("hbm_bytes", C.c_int64), ("precision", C.c_int32), ("block_size", C.c_int32),
                ("num_bermudans", C.c_int32), ("num_exercises", C.c_int32), ("num_continuation_dates", C.c_int32),
                ("max_exercises", C.c_int32), ("normals_read", C.c_int64)]


CCR_BF16 = 2
CCR_ACT_TANH = 0
CCR_ACT_SILU = 1


class MlpDescC(C.Structure):
    _fields_ = [("inputs", C.c_int32), ("hidden", C.c_int32), ("hidden_layers", C.c_int32), ("outputs", C.c_int32),
                ("activation", C.c_int32)]


class MlpTrainOptionsC(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("learning_rate", C.c_double), ("weight_decay", C.c_double),
                ("seed", C.c_uint32), ("precision", C.c_int32), ("warm", C.c_int32), ("batch", C.c_int32)]


class SurrogateOptionsC(C.Structure):
    _fields_ = [("train_paths", C.c_int32), ("hidden", C.c_int32), ("hidden_layers", C.c_int32), ("epochs", C.c_int32),
                ("learning_rate", C.c_double), ("weight_decay", C.c_double), ("seed", C.c_uint32),
                ("gate_threshold", C.c_double), ("batch", C.c_int32)]


class NestedImOutC(C.Structure):
    _fields_ = [("capacity", C.c_int32), ("num_dates", C.c_int32)] + [
        (k, C.POINTER(C.c_double)) for k in ("discounted", "mean", "band05", "band50", "band95")]


class SurrogateReportC(C.Structure):
    _fields_ = [("candidate_sets", C.c_int32), ("enabled_sets", C.c_int32), ("samples", C.c_int64),
                ("train_ms", C.c_double), ("loss", C.c_double), ("max_ee_error", C.c_double),
                ("networks", C.c_int32)]


def mlp_param_count(d: MlpDescC) -> int:
    sizes = [d.inputs] + [d.hidden] * d.hidden_layers + [d.outputs]
    return sum(sizes[i + 1] * sizes[i] + sizes[i + 1] for i in range(len(sizes) - 1))


PROFILE_FIELDS = ("times", "expected_value", "expected_exposure", "expected_negative_exposure",
                  "potential_future_exposure", "discounted_expected_exposure",
                  "discounted_expected_negative_exposure", "mean_deflator", "mean_survival", "cva_increments",
                  "dva_increments")


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


# ---------------------------------------------------------------- Python-side model objects

@dataclass
class Curve:
    times: Sequence[float]
    rates: Sequence[float]

    @staticmethod
    def flat(rate: float) -> "Curve":
        return Curve([1.0], [rate])


@dataclass
class CreditCurve:
    pillars: Sequence[float]
    hazards: Sequence[float]

    @staticmethod
    def flat(h: float) -> "CreditCurve":
        return CreditCurve([1.0], [h])


@dataclass
class HullWhite:
    curve: Curve
    a: float
    sigma: float


@dataclass
class Asset:
    spot: float
    vol: float
    carry: Curve
    foreign_rate: int = -1


@dataclass
class Cir:
    market: CreditCurve
    kappa: float
    theta: float
    xi: float
    y0: float
    substeps: int = 8


@dataclass
class Trade:
    kind: int
    notional: float
    fixed_rate: float = 0.0
    pay_fixed: bool = True
    schedule: Optional[Sequence[float]] = None
    float_schedule: Optional[Sequence[float]] = None
    currency: int = 0
    asset: int = 0
    maturity: float = 0.0
    exercise: Optional[Sequence[float]] = None
    cash_settled: bool = False
    position: float = 1.0
    degree: int = 3
    option_type: int = 0


@dataclass
class Collateral:
    threshold_counterparty: float = 0.0
    threshold_own: float = float("inf")
    mta: float = 0.0
    independent_amount: float = 0.0
    mpr: float = 10.0 / 250.0
    type: int = CCR_COLLATERAL_CASH
    haircut: float = 0.0
    bond_maturity: float = 0.0
    asset_index: int = 0


@dataclass
class NettingSet:
    trades: List[Trade]
    collateral: Optional[Collateral] = None
    counterparty_credit: int = -1
    own_credit: int = -1


@dataclass
class Model:
    rates: HullWhite
    foreign: List[HullWhite] = field(default_factory=list)
    assets: List[Asset] = field(default_factory=list)
    credits: List[Cir] = field(default_factory=list)
    correlation: Optional[np.ndarray] = None

    def num_factors(self) -> int:
        return 1 + len(self.foreign) + len(self.assets) + len(self.credits)

    def num_normals(self) -> int:
        return self.num_factors() + 1 + sum(c.substeps - 1 for c in self.credits) + len(self.foreign)


def swap_schedule(start: float, tenor: float, per_year: int) -> np.ndarray:
    """InterestRateSwap::vanilla schedule (src/instruments/interest_rate_swap.cpp:27-35)."""
    periods = int(np.floor(tenor * per_year + 0.5))  # std::lround for positive values
    s = np.array([start + i / per_year for i in range(periods + 1)], dtype=np.float64)
    s[-1] = start + tenor
    return s


def uniform_grid(horizon: float, steps: int) -> np.ndarray:
    """TimeGrid::uniform points before de-duplication (src/core/time_grid.cpp:22-27)."""
    return np.array([horizon * i / steps for i in range(steps + 1)], dtype=np.float64)


class Spec:
    """C structs for one workload: generator, simulation config, base grid, netting sets."""

    def __init__(self, model: Model, sets: List[NettingSet], base_grid: np.ndarray, num_paths: int,
                 seed: int = 42, antithetic: bool = True, pfe_quantile: float = 0.95, name: str = ""):
        self.model, self.sets, self.name = model, sets, name
        self.pfe_quantile = pfe_quantile
        self._keep = []
        self.base_grid = self._arr(base_grid)
        self.generator = self._generator(model)
        self.config = SimConfigC(num_paths, seed, 1 if antithetic else 0)
        self.sets_c = (NettingSetC * len(sets))()
        for i, ns in enumerate(sets):
            self.sets_c[i] = self._set(ns)

    @property
    def num_paths(self) -> int:
        return int(self.config.num_paths)

    def _arr(self, v) -> np.ndarray:
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        self._keep.append(a)
        return a

    def _curve(self, c: Curve) -> YieldCurveC:
        t, r = self._arr(c.times), self._arr(c.rates)
        return YieldCurveC(dptr(t), dptr(r), len(t))

    def _credit(self, c: CreditCurve) -> CreditCurveC:
        p, h = self._arr(c.pillars), self._arr(c.hazards)
        return CreditCurveC(dptr(p), dptr(h), len(p))

    def _hw(self, m: HullWhite) -> HullWhiteC:
        return HullWhiteC(self._curve(m.curve), m.a, m.sigma)

    def _generator(self, m: Model) -> GeneratorC:
        g = GeneratorC()
        g.rates = self._hw(m.rates)
        fr = (HullWhiteC * max(1, len(m.foreign)))()
        for i, f in enumerate(m.foreign):
            fr[i] = self._hw(f)
        ass = (AssetC * max(1, len(m.assets)))()
        for i, a in enumerate(m.assets):
            ass[i] = AssetC(a.spot, a.vol, self._curve(a.carry), a.foreign_rate)
        cr = (CirC * max(1, len(m.credits)))()
        for i, c in enumerate(m.credits):
            cr[i] = CirC(self._credit(c.market), c.kappa, c.theta, c.xi, c.y0, c.substeps)
        self._keep += [fr, ass, cr]
        g.num_foreign_rates, g.foreign_rates = len(m.foreign), fr
        g.num_assets, g.assets = len(m.assets), ass
        g.num_credits, g.credits = len(m.credits), cr
        if m.correlation is not None:
            corr = self._arr(np.asarray(m.correlation, dtype=np.float64).reshape(-1))
            g.correlation = dptr(corr)
        else:
            g.correlation = None
        return g

    def _trade(self, t: Trade) -> TradeC:
        c = TradeC()
        c.kind, c.pay_fixed, c.currency, c.asset = t.kind, 1 if t.pay_fixed else 0, t.currency, t.asset
        c.notional, c.fixed_rate, c.maturity = t.notional, t.fixed_rate, t.maturity
        if t.schedule is not None:
            s = self._arr(t.schedule)
            c.schedule, c.schedule_size = dptr(s), len(s)
        if t.float_schedule is not None:
            s = self._arr(t.float_schedule)
            c.float_schedule, c.float_schedule_size = dptr(s), len(s)
        if t.exercise is not None:
            e = self._arr(t.exercise)
            c.exercise_times, c.num_exercises = dptr(e), len(e)
        c.cash_settled, c.position, c.regression_degree = 1 if t.cash_settled else 0, t.position, t.degree
        c.option_type = t.option_type
        return c

    def _set(self, ns: NettingSet) -> NettingSetC:
        tr = (TradeC * len(ns.trades))()
        for i, t in enumerate(ns.trades):
            tr[i] = self._trade(t)
        self._keep.append(tr)
        c = NettingSetC()
        c.num_trades, c.trades = len(ns.trades), tr
        c.counterparty_credit, c.own_credit = ns.counterparty_credit, ns.own_credit
        if ns.collateral is not None:
            k = ns.collateral
            c.has_collateral = 1
            c.collateral = CollateralC(k.threshold_counterparty, k.threshold_own, k.mta, k.independent_amount,
                                       k.mpr, k.type, k.haircut, k.bond_maturity, k.asset_index)
        return c


class ProfileOut:
    """Caller-allocated ccr_exposure_out with numpy storage."""

    def __init__(self, capacity: int, num_paths: int = 0, rows: bool = False):
        self.capacity = capacity
        self.arrays = {n: np.full(capacity, np.nan) for n in PROFILE_FIELDS}
        self.rows = np.zeros(capacity * num_paths) if rows else None
        self.c = ExposureOutC()
        self.c.capacity = capacity
        for n in PROFILE_FIELDS:
            setattr(self.c, n, dptr(self.arrays[n]))
        self.c.exposure_rows = dptr(self.rows) if rows else None

    def get(self, name: str) -> np.ndarray:
        return self.arrays[name][: self.c.num_dates]

    def exposure_rows(self, num_paths: int) -> np.ndarray:
        return self.rows[: self.c.num_dates * num_paths].reshape(self.c.num_dates, num_paths)

    def as_dict(self):
        d = {n: self.get(n).copy() for n in PROFILE_FIELDS}
        d["num_dates"] = int(self.c.num_dates)
        return d


def make_outs(n_sets: int, capacity: int, num_paths: int = 0, rows: bool = False):
    outs = [ProfileOut(capacity, num_paths, rows) for _ in range(n_sets)]
    arr = (ExposureOutC * n_sets)()
    for i, o in enumerate(outs):
        arr[i] = o.c
    return outs, arr


def sync_outs(outs, arr):
    for i, o in enumerate(outs):
        o.c.num_dates = arr[i].num_dates
