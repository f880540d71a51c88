"""ccr_hip_execute_noalloc, the entry point that allocates nothing: a plan compiled with the
shared workspace executes with its per-execute buffers in the library's static scratch. Its
results are those of ccr_hip_execute on owned buffers, bit for bit, on consecutive calls of
different sizes (the scratch keeps the previous call's contents)."""
import numpy as np
import pytest

from zlatko_amd import abi, native
from zlatko_amd import workloads as W

pytestmark = pytest.mark.gpu


def _same(a, b, rows_paths=0):
    for f in abi.PROFILE_FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f
    if rows_paths:
        assert np.array_equal(a.exposure_rows(rows_paths), b.exposure_rows(rows_paths))


def test_noalloc_execute_matches_execute_on_consecutive_calls():
    # A larger plan first (three CSA sets in three currencies), then a smaller one holding
    # Bermudan swaptions (American Monte Carlo buffers in the scratch), then the first again.
    big = W.c2(num_paths=2048, num_sets=3, num_trades=30, weeks=26, horizon=0.5)
    small = W.small_reference_case(num_paths=600, dates=12, csa=True, fx=True, credits=True, two_leg=True,
                                   bermudans=2)
    ref = native.Engine()
    want_big = ref.run(big, rows=True)
    want_small = ref.run(small, rows=True)
    ref.close()
    e = native.Engine()
    e.set_shared_workspace(True)
    pb, ps = e.compile(big), e.compile(small)
    print("noalloc: big plan", pb.info().num_sets, "sets; small plan", ps.info().num_bermudans, "Bermudans", flush=True)
    got_big = pb.execute_noalloc(rows=True)
    for a, b in zip(got_big, want_big):
        _same(a, b, big.num_paths)
    got_small = ps.execute_noalloc(rows=True)
    _same(got_small[0], want_small[0], small.num_paths)
    again = pb.execute_noalloc()
    for a, b in zip(again, want_big):
        _same(a, b)
    # The ordinary execute of the same plan (context workspace) agrees as well.
    for a, b in zip(pb.execute(), want_big):
        _same(a, b)
    pb.close()
    ps.close()
    e.close()


def test_noalloc_execute_rejects_what_it_cannot_run():
    e = native.Engine()
    owned = e.compile(W.c1(num_paths=256, dates=12))
    with pytest.raises(native.CcrError) as x:
        owned.execute_noalloc()
    assert x.value.code == abi.CCR_ERR_INVALID_ARGUMENT
    owned.close()
    e.set_shared_workspace(True)
    large = e.compile(W.c1(num_paths=100000))  # ~97 MB of exposure rows alone
    with pytest.raises(native.CcrError) as x:
        large.execute_noalloc()
    assert x.value.code == abi.CCR_ERR_OUT_OF_RANGE and "64 MiB" in str(x.value)
    large.close()
    f32 = e.compile(W.c1(num_paths=256, dates=12), abi.CCR_FP32)
    with pytest.raises(native.CcrError) as x:
        f32.execute_noalloc(rows=True)
    assert x.value.code == abi.CCR_ERR_UNSUPPORTED
    assert f32.execute_noalloc()[0].c.num_dates > 0  # FP32 statistics without rows are fine
    f32.close()
    e.close()
