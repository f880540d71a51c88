"""Analysis specifications for the Spec/Value front door tests, in the reference's spec format
(src/app/analyses.cpp:85-290): two counterparties, one FX asset, correlation order [rates,
EURUSD, A, B], the market of the reference's own app tests (tests/test_app.cpp:24-56)."""
import copy


def credit(name, spread):
    return {"name": name, "recovery": 0.4,
            "quotes": [{"maturity": 1, "spread": spread}, {"maturity": 5, "spread": spread * 1.4}],
            "cir": {"kappa": 0.4, "theta": 0.02, "xi": 0.1, "y0": spread * 1.5, "substeps": 4}}


def swap(id_, notional, tenor, direction, fixed=0.034):
    return {"type": "swap", "id": id_, "notional": notional, "fixedRate": fixed, "start": 0, "tenor": tenor,
            "freq": 2, "direction": direction}


def forward(id_, notional, maturity):
    return {"type": "forward", "id": id_, "asset": "EURUSD", "notional": notional, "strike": 1.12,
            "maturity": maturity}


def netting_set(id_, cpty, trades, csa=False):
    return {"id": id_, "counterparty": cpty, "trades": trades,
            "csa": {"enabled": csa, "thresholdCounterparty": 1e5, "thresholdOwn": 1e5, "mta": 1e4, "mpr": 10.0 / 250.0}}


def base_spec(paths=1024):
    return {
        "domestic": {"times": [1, 5, 10], "rates": [0.03, 0.033, 0.035]},
        "hw": {"a": 0.04, "sigma": 0.009},
        "assets": [{"name": "EURUSD", "spot": 1.1, "vol": 0.1, "carry": {"flat": 0.02}}],
        "credits": [credit("A", 0.008), credit("B", 0.015)],
        "correlation": [[1, 0.2, 0.1, 0], [0.2, 1, 0, 0.3], [0.1, 0, 1, 0.5], [0, 0.3, 0.5, 1]],
        "sim": {"numPaths": paths, "seed": 7, "antithetic": True},
        "grid": {"type": "standard", "horizon": 5},
        "own": {"recovery": 0.4, "quotes": [{"maturity": 5, "spread": 0.006}]},
    }


def single_spec(analysis, paths=1024, trades=None, csa=False):
    """One netting set against counterparty A (its credit only, the matching correlation block)."""
    s = base_spec(paths)
    s["credits"] = [credit("A", 0.008)]
    s["correlation"] = [[1, 0.2, 0.1], [0.2, 1, 0], [0.1, 0, 1]]
    s["trades"] = trades if trades is not None else [swap("s", 1e7, 5, "payer"), forward("f", -5e6, 3)]
    if csa:
        s["csa"] = {"thresholdCounterparty": 1e5, "thresholdOwn": 1e5, "mta": 1e4, "mpr": 10.0 / 250.0}
    s["analysis"] = analysis
    return s


def book(spec, sets, analysis):
    s = copy.deepcopy(spec)
    s["scope"] = "book"
    s["analysis"] = analysis
    s["book"] = {"name": "test book", "nettingSets": sets}
    return s
