"""Profile comparison with scale-relative errors.

Errors are measured per field as max |a - b| over reporting dates divided by the field's scale:
its own max |b| for non-negative statistics (EE, EE*, PFE, E[D], E[Q], CVA/DVA increments), and
max(EE) + max(|ENE|) for the signed means E[V], ENE, ENE*, which can cancel to ~0 on at-par books.

pointwise_errors measures each date against its own value instead: max over dates of
|a_i - b_i| / max(|b_i|, POINTWISE_FLOOR x scale). The floor (1e-4 of the field's scale) keeps
dates whose statistic is ~0 -- at-par means, exposures before the first margin call -- from
demanding better than ~1e-14 of the scale in absolute terms, the rounding of a reordered sum of
terms of that size; every other date is held to the relative bar on its own value.
"""
from __future__ import annotations

import numpy as np

from zlatko_amd import abi

SIGNED = ("expected_value", "expected_negative_exposure", "discounted_expected_negative_exposure")
DISCOUNTED_SIGNED = ("discounted_expected_negative_exposure",)


def field_errors(a, b) -> dict:
    """a, b: ProfileOut (or dicts of arrays); returns {field: scale-relative max error}."""
    get = (lambda o, f: o.get(f)) if hasattr(a, "get") else (lambda o, f: np.asarray(o[f]))
    out = {}
    ee = np.abs(get(b, "expected_exposure")).max() if len(get(b, "expected_exposure")) else 0.0
    ene = np.abs(get(b, "expected_negative_exposure")).max() if len(get(b, "expected_negative_exposure")) else 0.0
    dee = np.abs(get(b, "discounted_expected_exposure")).max()
    dene = np.abs(get(b, "discounted_expected_negative_exposure")).max()
    for f in abi.PROFILE_FIELDS:
        x, y = np.asarray(get(a, f), dtype=np.float64), np.asarray(get(b, f), dtype=np.float64)
        assert x.shape == y.shape, (f, x.shape, y.shape)
        if f == "times":
            out[f] = 0.0 if np.array_equal(x, y) else float("inf")
            continue
        if f in DISCOUNTED_SIGNED:
            scale = dee + dene
        elif f in SIGNED:
            scale = ee + ene
        else:
            scale = np.abs(y).max()
        d = np.abs(x - y).max() if len(x) else 0.0
        out[f] = 0.0 if d == 0.0 else (d / scale if scale > 0 else float("inf"))
    return out


POINTWISE_FLOOR = 1e-4


def pointwise_errors(a, b) -> dict:
    """{field: max over dates of |a_i - b_i| / max(|b_i|, POINTWISE_FLOOR x the field's scale)}."""
    get = (lambda o, f: o.get(f)) if hasattr(a, "get") else (lambda o, f: np.asarray(o[f]))
    ee = np.abs(get(b, "expected_exposure")).max() if len(get(b, "expected_exposure")) else 0.0
    ene = np.abs(get(b, "expected_negative_exposure")).max() if len(get(b, "expected_negative_exposure")) else 0.0
    dee = np.abs(get(b, "discounted_expected_exposure")).max() if len(get(b, "expected_exposure")) else 0.0
    dene = np.abs(get(b, "discounted_expected_negative_exposure")).max() if len(get(b, "expected_exposure")) else 0.0
    out = {}
    for f in abi.PROFILE_FIELDS:
        if f == "times":
            continue
        x, y = np.asarray(get(a, f), dtype=np.float64), np.asarray(get(b, f), dtype=np.float64)
        if not len(x):
            out[f] = 0.0
            continue
        if f in DISCOUNTED_SIGNED:
            scale = dee + dene
        elif f in SIGNED:
            scale = ee + ene
        else:
            scale = np.abs(y).max()
        d = np.abs(x - y)
        den = np.maximum(np.abs(y), POINTWISE_FLOOR * scale)
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(d == 0.0, 0.0, d / np.where(den > 0, den, np.inf))
        r[(d > 0) & (den == 0)] = np.inf
        out[f] = float(r.max())
    return out


def worst(errs: dict) -> float:
    return max(errs.values())
