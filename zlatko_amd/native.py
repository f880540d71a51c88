"""ctypes binding of libccr_hip.so (include/ccr_hip.h).

The HIP library is the only compute path: if it is missing, importing this module's
entry points raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import json
import re
from pathlib import Path
from typing import List, Optional

import numpy as np

from . import abi

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libccr_hip.so"
HEADER = HERE.parent / "include" / "ccr_hip.h"
APP_HEADER = HERE.parent / "include" / "ccr_hip_app.h"

_lib = None


class CcrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ccr_hip error {code}: {message}")
        self.code = code


def lib():
    """Loads libccr_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with zlatko_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        I32, I64, D, P = C.c_int32, C.c_int64, C.c_double, C.c_void_p
        L.ccr_hip_abi_version.restype = I32
        L.ccr_hip_create.argtypes = [P, P]
        L.ccr_hip_create.restype = I32
        L.ccr_hip_destroy.argtypes = [P]
        L.ccr_hip_last_error.argtypes = [P]
        L.ccr_hip_last_error.restype = C.c_char_p
        L.ccr_hip_set_fault.argtypes = [P, I32]
        L.ccr_hip_set_shared_workspace.argtypes = [P, I32]
        L.ccr_hip_set_initial_margin.argtypes = [P, D, D]
        L.ccr_hip_set_nested_initial_margin.argtypes = [P, I32, D, D]
        L.ccr_hip_nested_initial_margin.argtypes = [P, P, I32, P]
        L.ccr_hip_initial_margin.argtypes = [P, P, I32, P, I32, P]
        L.ccr_hip_set_trade_statistics.argtypes = [P, I32]
        L.ccr_hip_set_book.argtypes = [P, I32]
        L.ccr_hip_set_book_grids.argtypes = [P, P, I32, P, I32]
        L.ccr_hip_set_book_group.argtypes = [P, I32]
        L.ccr_hip_set_regression.argtypes = [P, I32, I32]
        L.ccr_hip_book_grids.argtypes = [P, I32, P, I32, P, I32, P, P, I32, P]
        L.ccr_hip_set_graph.argtypes = [P, I32]
        L.ccr_book_totals.argtypes = [P, P, I32, I32, P]
        L.ccr_allocate_exposure_xva.argtypes = [P, P, I32, I32, P, D, P, P, P]
        L.ccr_hip_trade_statistics.argtypes = [P, P, I32, I32, P]
        L.ccr_xva_defaults.argtypes = [P]
        L.ccr_xva_defaults.restype = None
        L.ccr_exposure_xva.argtypes = [P, P, I32, P, D, P, P, P]
        L.ccr_exposure_xva.restype = I32
        L.ccr_funding_benefit.argtypes = [P, P, I32, P, P, D]
        L.ccr_funding_benefit.restype = D
        L.ccr_margin_valuation_adjustment.argtypes = [P, P, I32, P, P, D]
        L.ccr_margin_valuation_adjustment.restype = D
        L.ccr_xva.argtypes = [P, P, P, I32, P, P, D, P, D, P, P]
        L.ccr_xva.restype = I32
        L.ccr_hip_reporting_grid.argtypes = [P, I32, P, P, I32, P]
        L.ccr_hip_normals.argtypes = [C.c_uint64, C.c_int64, P]
        L.ccr_hip_simulation_grid.argtypes = [P, I32, P, P, I32, P]
        L.ccr_hip_compile.argtypes = [P, P, P, P, I32, P, I32, D, I32, P]
        L.ccr_hip_compile.restype = I32
        L.ccr_hip_plan_info_get.argtypes = [P, P]
        L.ccr_hip_execute.argtypes = [P, P, P]
        L.ccr_hip_execute.restype = I32
        L.ccr_hip_execute_async.argtypes = [P, P]
        L.ccr_hip_execute_async.restype = I32
        L.ccr_hip_execute_noalloc.argtypes = [P, P, P]
        L.ccr_hip_execute_noalloc.restype = I32
        L.ccr_hip_synchronize.argtypes = [P]
        L.ccr_hip_synchronize.restype = I32
        L.ccr_hip_last_timing.argtypes = [P, P]
        L.ccr_hip_plan_destroy.argtypes = [P]
        L.ccr_hip_run_exposure.argtypes = [P, P, P, P, I32, P, I32, D, I32, P]
        L.ccr_hip_run_exposure.restype = I32
        L.ccr_unilateral_cva.restype = D
        L.ccr_unilateral_cva.argtypes = [P, P, I32, P, D]
        L.ccr_bilateral_cva.argtypes = [P, P, P, I32, P, D, P, D, P, P]
        L.ccr_hip_mlp_train.restype = I32
        L.ccr_hip_mlp_train.argtypes = [P, P, P, P, P, I64, P, P, P, P]
        L.ccr_hip_mlp_gradient.restype = I32
        L.ccr_hip_mlp_gradient.argtypes = [P, P, I32, P, P, P, I64, P, P]
        L.ccr_hip_mlp_predict.restype = I32
        L.ccr_hip_mlp_predict.argtypes = [P, P, I32, P, P, P, I64, P]
        L.ccr_hip_surrogate_train.restype = I32
        L.ccr_hip_surrogate_train.argtypes = [P, P, P, P, P, P]
        L.ccr_hip_surrogate_set_info.restype = I32
        L.ccr_hip_surrogate_set_info.argtypes = [P, I32, P, P]
        L.ccr_hip_execute_surrogate.restype = I32
        L.ccr_hip_execute_surrogate.argtypes = [P, P, P, P]
        L.ccr_hip_execute_surrogate_async.restype = I32
        L.ccr_hip_execute_surrogate_async.argtypes = [P, P, P]
        L.ccr_hip_surrogate_destroy.argtypes = [P]
        L.ccr_hip_autotune.restype = I32
        L.ccr_hip_autotune.argtypes = [P, P, P, P, I32, P, I32, I32, I64, P, P]
        L.ccr_hip_autotune_book.restype = I32
        L.ccr_hip_autotune_book.argtypes = [P, P, P, P, I32, P, I32, I32, I64, P, P, P]
        L.ccr_hip_bermudan.restype = I32
        L.ccr_hip_bermudan.argtypes = [P, P, P, P, I32, P, I32, P, P]
        L.ccr_expected_positive_exposure.restype = D
        L.ccr_expected_positive_exposure.argtypes = [P, P, I32, D]
        L.ccr_effective_expected_exposure.argtypes = [P, I32, P]
        L.ccr_effective_expected_positive_exposure.restype = D
        L.ccr_effective_expected_positive_exposure.argtypes = [P, P, I32, D]
        L.ccr_hip_app_call.restype = I32
        L.ccr_hip_app_call.argtypes = [P, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.ccr_hip_app_free.restype = None
        L.ccr_hip_app_free.argtypes = [C.c_void_p]
        L.ccr_hip_app_functions.restype = C.c_char_p
        L.ccr_hip_app_functions.argtypes = []
        _lib = L
    return _lib


def header_symbols() -> List[str]:
    """Function names declared in include/ccr_hip.h and include/ccr_hip_app.h."""
    text = HEADER.read_text() + APP_HEADER.read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(ccr_\w+)\s*\(", text, re.M)))


def app_call(function: str, spec, engine: "Engine" = None) -> dict:
    """ccr_hip_app_call: the Spec/Value front door. spec is a JSON-serialisable dict (or JSON
    text); returns the result document as a dict ({"error": ...} for a failed analysis)."""
    text = spec if isinstance(spec, str) else json.dumps(spec)
    out = C.c_void_p()
    rc = lib().ccr_hip_app_call(engine.ctx if engine else None, function.encode(), text.encode(), C.byref(out))
    if rc:
        raise CcrError(rc, "ccr_hip_app_call")
    try:
        return json.loads(C.string_at(out.value).decode())
    finally:
        lib().ccr_hip_app_free(out)


def grid(spec: abi.Spec, i: int = 0, simulation: bool = False) -> np.ndarray:
    fn = lib().ccr_hip_simulation_grid if simulation else lib().ccr_hip_reporting_grid
    n = C.c_int32()
    rc = fn(abi.dptr(spec.base_grid), len(spec.base_grid), C.byref(spec.sets_c[i]), None, 0, C.byref(n))
    if rc:
        raise CcrError(rc, "grid")
    out = np.zeros(n.value)
    fn(abi.dptr(spec.base_grid), len(spec.base_grid), C.byref(spec.sets_c[i]), abi.dptr(out), n.value, C.byref(n))
    return out


def normals(seed: int, n: int) -> np.ndarray:
    """The replayed normal stream (ccr::NormalGenerator(seed), n draws), on the host's cores."""
    out = np.empty(n, np.float64)
    rc = lib().ccr_hip_normals(seed, n, out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"ccr_hip_normals failed ({rc})")
    return out


def book_grids(spec: abi.Spec):
    """ccr_hip_book_grids: (reporting, simulation) union grids of all of spec's netting sets."""
    nr, ns = C.c_int32(), C.c_int32()
    args = (abi.dptr(spec.base_grid), len(spec.base_grid), spec.sets_c, len(spec.sets))
    rc = lib().ccr_hip_book_grids(*args, None, 0, C.byref(nr), None, 0, C.byref(ns))
    if rc:
        raise CcrError(rc, "book grids")
    r, v = np.zeros(nr.value), np.zeros(ns.value)
    lib().ccr_hip_book_grids(*args, abi.dptr(r), len(r), C.byref(nr), abi.dptr(v), len(v), C.byref(ns))
    return r, v


class Plan:
    def __init__(self, engine: "Engine", handle: C.c_void_p, spec: abi.Spec):
        self.engine, self.handle, self.spec = engine, handle, spec

    def info(self) -> abi.HipPlanInfoC:
        i = abi.HipPlanInfoC()
        lib().ccr_hip_plan_info_get(self.handle, C.byref(i))
        return i

    def execute(self, rows: bool = False):
        info = self.info()
        outs, arr = abi.make_outs(len(self.spec.sets), info.max_reporting_dates, self.spec.num_paths, rows)
        self.engine._check(lib().ccr_hip_execute(self.engine.ctx, self.handle, arr))
        abi.sync_outs(outs, arr)
        return outs

    def execute_async(self):
        self.engine._check(lib().ccr_hip_execute_async(self.engine.ctx, self.handle))

    def execute_noalloc(self, rows: bool = False):
        """ccr_hip_execute_noalloc: the per-execute buffers in the library's static scratch."""
        info = self.info()
        outs, arr = abi.make_outs(len(self.spec.sets), info.max_reporting_dates, self.spec.num_paths, rows)
        self.engine._check(lib().ccr_hip_execute_noalloc(self.engine.ctx, self.handle, arr))
        abi.sync_outs(outs, arr)
        return outs

    def trade_statistics(self, s: int, trade: int) -> dict:
        """Per-trade statistics of trade `trade` of netting set `s` from the last execute: one array
        per name of abi.TRADE_STATS, one value per reporting date."""
        info = self.info()
        arrays = {n: np.zeros(max(1, info.max_reporting_dates)) for n in abi.TRADE_STATS}
        out = abi.TradeStatisticsC(info.max_reporting_dates, 0, *[abi.dptr(arrays[n]) for n in abi.TRADE_STATS])
        self.engine._check(lib().ccr_hip_trade_statistics(self.engine.ctx, self.handle, s, trade, C.byref(out)))
        return {n: a[:out.num_dates].copy() for n, a in arrays.items()}

    def initial_margin(self, s: int = 0) -> np.ndarray:
        """initialMarginProfile (src/xva/xva.cpp:59-76) of netting set `s` from the last execute,
        one value per simulation date."""
        info = self.info()
        out = np.zeros(max(1, info.max_sim_dates))
        n = C.c_int32()
        self.engine._check(lib().ccr_hip_initial_margin(self.engine.ctx, self.handle, s, abi.dptr(out), len(out),
                                                        C.byref(n)))
        return out[:n.value].copy()

    def nested_initial_margin(self, s: int = 0) -> dict:
        """Nested conditional initial margin of netting set `s` from the last execute, per reporting
        date: discounted (E[D IM]), mean (E[IM]), band05 / band50 / band95 (quantiles of IM)."""
        cap = max(1, self.info().max_reporting_dates)
        arrs = {k: np.zeros(cap) for k in ("discounted", "mean", "band05", "band50", "band95")}
        out = abi.NestedImOutC(cap, 0, *(abi.dptr(arrs[k]) for k in ("discounted", "mean", "band05", "band50",
                                                                         "band95")))
        self.engine._check(lib().ccr_hip_nested_initial_margin(self.engine.ctx, self.handle, s, C.byref(out)))
        return {k: v[:out.num_dates].copy() for k, v in arrs.items()}

    def execute_surrogate(self, surrogate: "Surrogate", rows: bool = False):
        info = self.info()
        outs, arr = abi.make_outs(len(self.spec.sets), info.max_reporting_dates, self.spec.num_paths, rows)
        self.engine._check(lib().ccr_hip_execute_surrogate(self.engine.ctx, self.handle, surrogate.handle, arr))
        abi.sync_outs(outs, arr)
        return outs

    def execute_surrogate_async(self, surrogate: "Surrogate"):
        self.engine._check(lib().ccr_hip_execute_surrogate_async(self.engine.ctx, self.handle, surrogate.handle))

    def close(self):
        if self.handle:
            lib().ccr_hip_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """A trained exposure surrogate (per netting set with Bermudans, gated on hold-out paths)."""

    def __init__(self, handle: C.c_void_p, report: abi.SurrogateReportC):
        self.handle, self.report = handle, report

    def set_info(self, s: int):
        en, err = C.c_int32(), C.c_double()
        lib().ccr_hip_surrogate_set_info(self.handle, s, C.byref(en), C.byref(err))
        return en.value, err.value

    def close(self):
        if self.handle:
            lib().ccr_hip_surrogate_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """A ccr_hip context on one device."""

    def __init__(self, device: int = 0, block_size: int = 0):
        self.ctx = C.c_void_p()
        opts = abi.HipOptionsC(device, abi.CCR_FP64, block_size)
        rc = lib().ccr_hip_create(C.byref(opts), C.byref(self.ctx))
        if rc:
            raise CcrError(rc, lib().ccr_hip_last_error(None).decode())

    def _check(self, rc: int):
        if rc:
            raise CcrError(rc, lib().ccr_hip_last_error(self.ctx).decode())

    def compile(self, spec: abi.Spec, precision: int = abi.CCR_FP64) -> Plan:
        h = C.c_void_p()
        self._check(lib().ccr_hip_compile(self.ctx, C.byref(spec.generator), C.byref(spec.config),
                                          abi.dptr(spec.base_grid), len(spec.base_grid), spec.sets_c,
                                          len(spec.sets), spec.pfe_quantile, precision, C.byref(h)))
        return Plan(self, h, spec)

    def run(self, spec: abi.Spec, precision: int = abi.CCR_FP64, rows: bool = False):
        cap = max(len(grid(spec, i)) for i in range(len(spec.sets)))
        outs, arr = abi.make_outs(len(spec.sets), cap, spec.num_paths, rows)
        self._check(lib().ccr_hip_run_exposure(self.ctx, C.byref(spec.generator), C.byref(spec.config),
                                               abi.dptr(spec.base_grid), len(spec.base_grid), spec.sets_c,
                                               len(spec.sets), spec.pfe_quantile, precision, arr))
        abi.sync_outs(outs, arr)
        return outs

    def timing(self) -> abi.HipTimingC:
        t = abi.HipTimingC()
        lib().ccr_hip_last_timing(self.ctx, C.byref(t))
        return t

    def synchronize(self):
        self._check(lib().ccr_hip_synchronize(self.ctx))

    def set_fault(self, kind: int):
        self._check(lib().ccr_hip_set_fault(self.ctx, kind))

    def set_shared_workspace(self, enable: bool):
        """Plans compiled from now on keep their per-execute buffers in one workspace of the
        context (valid until the next execute of any plan), so waves of netting sets fit in HBM."""
        self._check(lib().ccr_hip_set_shared_workspace(self.ctx, 1 if enable else 0))

    def set_graph(self, enable: bool):
        """Executes replay one captured HIP graph per plan from now on."""
        self._check(lib().ccr_hip_set_graph(self.ctx, 1 if enable else 0))

    def set_book(self, enable: bool):
        """Plans compiled from now on put every netting set on the book's common grids (pathwise
        identical scenarios, additive statistics)."""
        self._check(lib().ccr_hip_set_book(self.ctx, 1 if enable else 0))

    def set_book_group(self, sets_per_workgroup: int):
        self._check(lib().ccr_hip_set_book_group(self.ctx, sets_per_workgroup))

    def set_regression(self, basis: int, gram_precision: int = 0):
        """Bermudan regressions of plans compiled from now on: basis -1 (the state basis on shared
        per-date designs), BASIS_STATE or BASIS_SWAP_RATE_ANNUITY on the regression engine, its
        Grams in FP64 (gram_precision FP64) or on FP32 matrix cores (FP32)."""
        self._check(lib().ccr_hip_set_regression(self.ctx, basis, gram_precision))

    def set_book_grids(self, reporting, simulation):
        """Book plans compiled afterwards use these grids (None: their own union grids)."""
        if reporting is None:
            self._check(lib().ccr_hip_set_book_grids(self.ctx, None, 0, None, 0))
            return
        self._book_grids = (np.ascontiguousarray(reporting, np.float64), np.ascontiguousarray(simulation, np.float64))
        r, v = self._book_grids
        self._check(lib().ccr_hip_set_book_grids(self.ctx, abi.dptr(r), len(r), abi.dptr(v), len(v)))

    def set_trade_statistics(self, enable: bool):
        """Plans compiled from now on also produce per-trade allocation statistics (at most 16
        trades per netting set, no Bermudan swaptions)."""
        self._check(lib().ccr_hip_set_trade_statistics(self.ctx, 1 if enable else 0))

    def set_nested_initial_margin(self, inner_paths: int, quantile: float = 0.99,
                                  margin_period_of_risk: float = 10.0 / 250.0):
        """Plans compiled from now on also produce the nested conditional initial margin of every
        netting set (inner_paths 0 switches it off)."""
        self._check(lib().ccr_hip_set_nested_initial_margin(self.ctx, inner_paths, quantile, margin_period_of_risk))

    def set_initial_margin(self, quantile: float, margin_period_of_risk: float = 10.0 / 250.0):
        """Plans compiled from now on also produce the initial-margin profile of every netting set
        (quantile <= 0 switches it off)."""
        self._check(lib().ccr_hip_set_initial_margin(self.ctx, quantile, margin_period_of_risk))

    def autotune_book(self, spec: abi.Spec, precision: int = abi.CCR_FP64, sample_paths: int = 16384):
        """Picks the book kernel's workgroup size and netting sets per workgroup for this workload
        (book mode); returns (size, sets per workgroup, times in ms [4 sizes][4 groups], 0 = not run)."""
        block, group = C.c_int32(), C.c_int32()
        times = np.zeros(16)
        self._check(lib().ccr_hip_autotune_book(self.ctx, C.byref(spec.generator), C.byref(spec.config),
                                                abi.dptr(spec.base_grid), len(spec.base_grid), spec.sets_c,
                                                len(spec.sets), precision, sample_paths, C.byref(block),
                                                C.byref(group), abi.dptr(times)))
        return block.value, group.value, times.reshape(4, 4)

    def autotune(self, spec: abi.Spec, precision: int = abi.CCR_FP64, sample_paths: int = 16384):
        """Picks the path kernel's workgroup size for this workload; returns (size, times in ms)."""
        block = C.c_int32()
        times = np.zeros(4)
        self._check(lib().ccr_hip_autotune(self.ctx, C.byref(spec.generator), C.byref(spec.config),
                                           abi.dptr(spec.base_grid), len(spec.base_grid), spec.sets_c,
                                           len(spec.sets), precision, sample_paths, C.byref(block), abi.dptr(times)))
        return block.value, times

    def bermudan(self, spec: abi.Spec, set_index: int, trade_index: int, grid: np.ndarray,
                 precision: int = abi.CCR_FP64):
        """BermudanSwaption::valueWithExercise on the GPU: (price, exercise index per path, -1 never)."""
        grid = np.ascontiguousarray(grid, np.float64)
        price = C.c_double()
        ex = np.zeros(spec.num_paths, np.int32)
        trade = spec.sets_c[set_index].trades[trade_index]
        self._check(lib().ccr_hip_bermudan(self.ctx, C.byref(spec.generator), C.byref(spec.config), abi.dptr(grid),
                                           len(grid), C.byref(trade), precision, C.byref(price),
                                           ex.ctypes.data_as(C.c_void_p)))
        return price.value, ex

    # ------------------------------------------------------------ MLP surrogate

    def surrogate_train(self, train: Plan, holdout: Plan, train_paths: int = 2000, hidden: int = 128,
                        hidden_layers: int = 3, epochs: int = 300, lr: float = 0.003, wd: float = 1e-6,
                        seed: int = 1, gate: float = 0.005, batch: int = 0) -> Surrogate:
        o = abi.SurrogateOptionsC(train_paths, hidden, hidden_layers, epochs, lr, wd, seed, gate, batch)
        h, rep = C.c_void_p(), abi.SurrogateReportC()
        self._check(lib().ccr_hip_surrogate_train(self.ctx, train.handle, holdout.handle, C.byref(o), C.byref(h),
                                                  C.byref(rep)))
        return Surrogate(h, rep)

    def mlp_train(self, desc: abi.MlpDescC, X: np.ndarray, Y: np.ndarray, epochs: int, lr: float = 0.01,
                  wd: float = 1e-6, seed: int = 1, precision: int = abi.CCR_FP32, params=None, stdz=None,
                  batch: int = 0):
        """fitMlp on the GPU; returns (params, stdz, loss, train_ms). params/stdz given: warm start."""
        X = np.ascontiguousarray(X, np.float64)
        Y = np.ascontiguousarray(Y, np.float64)
        warm = params is not None
        params = np.zeros(abi.mlp_param_count(desc)) if params is None else np.array(params, np.float64)
        stdz = np.zeros(2 * (desc.inputs + desc.outputs)) if stdz is None else np.array(stdz, np.float64)
        o = abi.MlpTrainOptionsC(epochs, lr, wd, seed, precision, 1 if warm else 0, batch)
        loss, ms = C.c_double(), C.c_double()
        self._check(lib().ccr_hip_mlp_train(self.ctx, C.byref(desc), C.byref(o), abi.dptr(X), abi.dptr(Y), X.shape[0],
                                            abi.dptr(params), abi.dptr(stdz), C.byref(loss), C.byref(ms)))
        return params, stdz, loss.value, ms.value

    def mlp_gradient(self, desc: abi.MlpDescC, precision: int, params: np.ndarray, xs: np.ndarray, ys: np.ndarray):
        xs = np.ascontiguousarray(xs, np.float64)
        ys = np.ascontiguousarray(ys, np.float64)
        params = np.ascontiguousarray(params, np.float64)
        g = np.zeros_like(params)
        loss = C.c_double()
        self._check(lib().ccr_hip_mlp_gradient(self.ctx, C.byref(desc), precision, abi.dptr(params), abi.dptr(xs),
                                               abi.dptr(ys), xs.shape[0], abi.dptr(g), C.byref(loss)))
        return loss.value, g

    def mlp_predict(self, desc: abi.MlpDescC, precision: int, params: np.ndarray, stdz: np.ndarray, X: np.ndarray):
        X = np.ascontiguousarray(X, np.float64)
        Y = np.zeros((X.shape[0], desc.outputs))
        self._check(lib().ccr_hip_mlp_predict(self.ctx, C.byref(desc), precision,
                                              abi.dptr(np.ascontiguousarray(params, np.float64)),
                                              abi.dptr(np.ascontiguousarray(stdz, np.float64)), abi.dptr(X),
                                              X.shape[0], abi.dptr(Y)))
        return Y

    def close(self):
        if self.ctx:
            lib().ccr_hip_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
