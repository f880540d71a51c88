"""The Spec/Value front door on the GPU (ccr_hip_app_call "hipAnalyse"): each analysis against
the FP64 oracle run on the same parsed inputs, the identities of its result shape, and the
reference's own app tests of the portfolio book (tests/test_app.cpp:66-132: netting sets share
their scenarios, a book of one set is the set, CVA and XVA add up by set and counterparty)."""
import math

import numpy as np
import pytest

from tests import app_specs as A
from tests import oracle_lib as OL
from zlatko_amd import abi, native

pytestmark = pytest.mark.gpu

FIELDS = {"times": "times", "expectedValue": "expected_value", "ee": "expected_exposure",
          "ene": "expected_negative_exposure", "pfe": "potential_future_exposure",
          "discountedEe": "discounted_expected_exposure", "discountedEne": "discounted_expected_negative_exposure"}


@pytest.fixture(scope="module")
def engine():
    e = native.Engine()
    yield e
    e.close()


def analyse(engine, spec):
    r = native.app_call("hipAnalyse", spec, engine)
    assert "error" not in r and "unsupported" not in r, r
    return r


def engine_spec(spec):
    """The abi.Spec the front door compiles for a single-set spec, rebuilt from the spec and the
    parse it reports (hipJobs: bootstrapped curves, schedules, par rates)."""
    j = native.app_call("hipJobs", spec)
    dom = spec.get("domestic", {"flat": 0.03})
    curve = abi.Curve([1.0], [dom["flat"]]) if "flat" in dom else abi.Curve(dom["times"], dom["rates"])
    hw = spec.get("hw", {})
    model = abi.Model(abi.HullWhite(curve, hw.get("a", 0.05), hw.get("sigma", 0.01)))
    names = []
    for a in spec.get("assets", []):
        carry = a.get("carry", {"flat": 0.01})
        model.assets.append(abi.Asset(a.get("spot", 1.0), a.get("vol", 0.1),
                                      abi.Curve([1.0], [carry["flat"]]) if "flat" in carry
                                      else abi.Curve(carry["times"], carry["rates"])))
        names.append(a.get("name", "FX"))
    for c, cs in zip(j["credits"], spec.get("credits", [])):
        cir = cs.get("cir", {})
        model.credits.append(abi.Cir(abi.CreditCurve(c["pillars"], c["hazards"]), cir.get("kappa", 0.4),
                                     cir.get("theta", 0.02), cir.get("xi", 0.08), cir.get("y0", 0.01),
                                     cir.get("substeps", 4)))
    if "correlation" in spec:
        model.correlation = np.asarray(spec["correlation"], dtype=np.float64)
    trades = []
    for t, ts in zip(j["trades"], spec["trades"]):
        kind = ts.get("type", "swap")
        if kind == "swap":
            trades.append(abi.Trade(abi.CCR_TRADE_SWAP, t["notional"], t["fixedRate"], t["direction"] == "payer",
                                    schedule=t["schedule"]))
        elif kind == "forward":
            trades.append(abi.Trade(abi.CCR_TRADE_ASSET_FORWARD, t["notional"], t["fixedRate"],
                                    asset=names.index(ts.get("asset", "FX")), maturity=t["maturity"]))
        else:
            raise AssertionError(kind)
    csa = None
    if "csa" in spec and spec["csa"].get("enabled", True):
        c = spec["csa"]
        own = c.get("thresholdOwn", math.inf)
        csa = abi.Collateral(c.get("thresholdCounterparty", 0.0), own if own >= 0 else math.inf, c.get("mta", 0.0),
                             c.get("independentAmount", 0.0), c.get("mpr", 10.0 / 250.0))
    ns = abi.NettingSet(trades, csa, counterparty_credit=spec.get("counterparty", 0) if model.credits else -1)
    sim = spec.get("sim", {})
    return abi.Spec(model, [ns], np.asarray(j["baseGrid"]), sim.get("numPaths", 2000), sim.get("seed", 42),
                    sim.get("antithetic", True), spec.get("pfeQuantile", 0.95), "app")


def profile_of(result_profile):
    return {FIELDS[k]: np.asarray([np.nan if x is None else x for x in result_profile[k]]) for k in FIELDS}


@pytest.mark.parametrize("csa", [False, True])
def test_exposure_matches_the_oracle(engine, csa):
    spec = A.single_spec("exposure", paths=512, csa=csa)
    r = analyse(engine, spec)
    ref = OL.run_oracle(engine_spec(spec), threads=8)[0].as_dict()
    got = profile_of(r["profile"])
    assert np.array_equal(got["times"], ref["times"])
    # Scale-relative errors: max |got - oracle| over the dates / (max EE + max |ENE|) of the oracle.
    scale = np.abs(ref["expected_exposure"]).max() + np.abs(ref["expected_negative_exposure"]).max()
    errs = {f: float(np.abs(got[f] - ref[f]).max() / scale) for f in got if f != "times"}
    print("app exposure vs oracle", csa, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-10
    assert r["numPaths"] == 512 and r["reportingDates"] == len(ref["times"]) and r["collateralised"] == csa
    p = r["profile"]
    assert p["effectiveEe"] == list(np.maximum.accumulate(p["ee"]))
    assert p["maxPfe"] == max(p["pfe"]) and p["pfeQuantile"] == 0.95


def test_cva_analysis(engine):
    spec = A.single_spec("cva", paths=512)
    r = analyse(engine, spec)
    oracle = OL.run_oracle(engine_spec(spec), threads=8)[0].as_dict()
    rec = 0.4
    want = (1 - rec) * np.sum(oracle["cva_increments"][1:])
    assert r["pathwiseCva"] == pytest.approx(want, rel=1e-10)
    assert r["pathwiseExcess"] == pytest.approx(r["pathwiseCva"] / r["cva"] - 1, rel=1e-12)
    assert r["cumulativeCva"][-1] == pytest.approx(r["cva"], rel=1e-12)
    assert sum(r["termStructure"]) == pytest.approx(r["cva"], rel=1e-12)
    # CVA is linear in the loss given default.
    for R, c in zip(r["recoveryGrid"], r["cvaByRecovery"]):
        assert c == pytest.approx(r["cva"] * (1 - R) / (1 - rec), rel=1e-12)
    b = r["bilateral"]
    assert b["total"] == pytest.approx(b["cva"] - b["dva"], rel=1e-12) and b["dva"] > 0
    assert r["spreadMaturity"] == 5 and r["spreadNotional"] == 1e7 and r["runningSpread"] > 0
    if OL.ref() is not None and hasattr(OL.ref(), "ccr_ref_risky_annuity"):
        j = native.app_call("hipJobs", spec)
        ann = OL.ref_risky_annuity(spec["domestic"]["times"], spec["domestic"]["rates"], j["credits"][0]["pillars"],
                                   j["credits"][0]["hazards"], 5.0)
        assert r["runningSpread"] == pytest.approx(r["cva"] / (1e7 * ann), rel=1e-12)


def test_xva_analysis_with_nested_initial_margin(engine):
    spec = A.single_spec("xva", paths=512)
    spec["imInnerPaths"] = 64
    spec["imQuantile"] = 0.9
    r = analyse(engine, spec)
    assert r["imMethod"] == "nested" and r["identityHolds"]
    x = r["xva"]
    assert x["total"] == pytest.approx(x["cva"] - x["dva"] + x["fca"] - x["fba"] + x["mva"] + x["kva"], rel=1e-12)
    assert x["mva"] > 0 and len(r["initialMarginBands"]) == 3
    assert r["cumulativeCva"][-1] == pytest.approx(x["cva"], rel=1e-12)
    assert r["fvaBySpread"][0] == 0.0 and r["kvaByHurdle"][0] == 0.0
    # The Euler allocation of CVA, FCA and KVA adds up (no CSA: the trades' values net exactly).
    assert len(r["trades"]) == 2
    for k in ("cva", "fca", "kva"):
        assert sum(t[k] for t in r["trades"]) == pytest.approx(x[k], rel=1e-9), k
    spec["imMethod"] = "proxy"
    p = analyse(engine, spec)
    assert p["imMethod"] == "proxy" and p["initialMargin"] == p["initialMarginProxy"]
    assert p["initialMargin"][0] == 0.0 and max(p["initialMargin"]) > 0


def test_allocation_collateral_and_wrong_way_risk(engine):
    r = analyse(engine, A.single_spec("allocation", paths=512))
    assert r["addsUp"] and len(r["trades"]) == 2
    assert r["marginalCvaSum"] == pytest.approx(r["totalCva"], rel=1e-9)
    c = analyse(engine, A.single_spec("collateral", paths=512, csa=True))
    names = [v["name"] for v in c["variants"]]
    assert names[:2] == ["no CSA", "specified CSA"] and len(names) == 7 and names[-1].startswith("EURUSD")
    assert c["variants"][0]["eepeVsNoCsa"] == 1.0
    assert all(0.0 < v["eepeVsNoCsa"] < 1.0 for v in c["variants"][1:])  # every CSA cuts the exposure here
    assert not c["variants"][0]["csa"]["enabled"] and c["variants"][4]["csa"]["oneWay"]
    w = analyse(engine, dict(A.single_spec("wrongWayRisk", paths=512), rhos=[-0.5, 0.0, 0.5]))
    assert [x["rho"] for x in w["results"]] == [-0.5, 0.0, 0.5] and w["driver"] == 1
    for x in w["results"]:
        assert "error" not in x and x["multiplier"] > 0 and x["conditionalDiscountedEe"][0] is None


def test_book_netting_sets_share_their_scenarios(engine):
    sets = [A.netting_set("x", "A", [A.swap("s", 1e7, 5, "payer")]),
            A.netting_set("y", "B", [A.swap("s", 1e7, 5, "payer"), A.forward("f", 5e6, 3)]),
            A.netting_set("z", "A", [A.forward("f", 5e6, 3)])]
    r = analyse(engine, A.book(A.base_spec(), sets, "exposure"))
    assert r["scope"] == "book" and len(r["nettingSets"]) == 3
    prof = [s["result"]["profile"] for s in r["nettingSets"]]
    evx, evy, evz = (np.asarray(p["expectedValue"]) for p in prof)
    eex, eey, eez = (np.asarray(p["ee"]) for p in prof)
    assert len(evx) == len(evy) == len(evz) > 10
    # Shared scenarios: the book's values add up to FP64 rounding, not Monte Carlo error.
    assert np.abs(evy - (evx + evz)).max() <= 1e-9 * 1.5e7
    assert np.all(eey <= eex + eez + 1e-9 * 1.5e7)
    total = np.asarray(r["total"]["profile"]["ee"])
    assert np.allclose(total, eex + eey + eez, rtol=1e-12, atol=0)


def test_book_of_one_netting_set_is_the_netting_set(engine):
    trades = [A.swap("s", 1e7, 5, "payer"), A.forward("f", -5e6, 3)]
    single = A.single_spec("cva", paths=1024, trades=trades)
    s = analyse(engine, single)
    r = analyse(engine, A.book(A.base_spec(), [A.netting_set("only", "A", trades)], "cva"))
    assert r["total"]["cva"] == pytest.approx(s["cva"], rel=1e-9)
    assert r["total"]["pathwiseCva"] == pytest.approx(s["pathwiseCva"], rel=1e-9)


def test_book_xva_and_cva_add_up_by_netting_set_and_counterparty(engine):
    sets = [A.netting_set("x", "A", [A.swap("s1", 1e7, 5, "payer"), A.forward("f", 5e6, 3)]),
            A.netting_set("y", "B", [A.swap("s2", 6e6, 4, "receiver")], True),
            A.netting_set("z", "A", [A.forward("g", -3e6, 2)])]
    x = analyse(engine, A.book(A.base_spec(), sets, "xva"))
    total = sum(s["summary"]["total"] for s in x["nettingSets"])
    cva = sum(s["summary"]["cva"] for s in x["nettingSets"])
    assert x["total"]["xva"]["total"] == pytest.approx(total, rel=1e-9)
    by = x["total"]["byCounterparty"]
    assert [b["name"] for b in by] == ["A", "B"] and by[0]["nettingSets"] == ["x", "z"]
    assert by[0]["cva"] + by[1]["cva"] == pytest.approx(cva, rel=1e-9)
    assert len(x["total"]["trades"]) == 4
    c = analyse(engine, A.book(A.base_spec(), sets, "cva"))
    assert c["total"]["cva"] == pytest.approx(sum(s["summary"]["cva"] for s in c["nettingSets"]), rel=1e-12)
    a = analyse(engine, A.book(A.base_spec(), sets, "allocation"))
    assert len(a["total"]["trades"]) == 4
    assert "error" in native.app_call("hipAnalyse", A.book(A.base_spec(), sets, "models"), engine)
    bad = A.book(A.base_spec(), [A.netting_set("x", "C", [A.swap("s", 1e6, 5, "payer")])], "cva")
    assert "unknown counterparty 'C'" in native.app_call("hipAnalyse", bad, engine)["error"]
    spec = dict(A.base_spec(), rhos=[-0.5, 0.5])
    for analysis in ("collateral", "wrongWayRisk"):
        b = analyse(engine, A.book(spec, sets, analysis))
        assert len(b["nettingSets"]) == 3 and b["counterparties"] == ["A", "B"]
    col = analyse(engine, A.book(spec, sets, "collateral"))
    assert [v["name"] for v in col["total"]["variants"]][:2] == ["no CSA", "specified CSA"]
    assert col["total"]["variants"][0]["eepeVsNoCsa"] == 1.0
    wwr = analyse(engine, A.book(spec, sets, "wrongWayRisk"))
    for k, t in enumerate(wwr["total"]["results"]):
        assert t["pathwiseCva"] == pytest.approx(
            sum(s["result"]["results"][k]["pathwiseCva"] for s in wwr["nettingSets"]), rel=1e-12)
    assert "unsupported" in native.app_call("hipAnalyse", A.book(spec, sets, "hedging"), engine)
