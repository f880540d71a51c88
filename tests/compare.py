"""Profile comparison with scale-relative errors.

Errors are measured per field as max |a - b| over reporting dates divided by the field's scale:
its own max |b| for non-negative statistics (EE, EE*, PFE, E[D], E[Q], CVA/DVA increments), and
max(EE) + max(|ENE|) for the signed means E[V], ENE, ENE*, which can cancel to ~0 on at-par books.
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


def worst(errs: dict) -> float:
    return max(errs.values())
