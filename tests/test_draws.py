"""The replayed normal stream generated on the host's cores (ccr_hip_normals, the stream every
plan uploads) against the sequential ccr::NormalGenerator (include/ccr/core/random.hpp:13-26) as
the oracle restates it, and against the reference's golden draws: bit for bit, across batch and
chunk boundaries, for odd lengths (a normal held back by the polar method) and several seeds."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib
from zlatko_amd import native

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def test_golden_draws_of_the_reference():
    assert np.array_equal(native.normals(42, 64), np.array(GOLDEN["normals_seed42"]))


@pytest.mark.parametrize("seed,n", [(42, 0), (42, 1), (42, 2), (7, 3), (7, 102_899), (2024, 1_000_001),
                                    (123456789, 9_000_003), (2**64 - 1, 250_000)])
def test_parallel_stream_is_the_sequential_stream(seed, n):
    assert np.array_equal(native.normals(seed, n), oracle_lib.normals(seed, n))


def test_stream_prefixes_agree():
    """A shorter request is a prefix of a longer one (plans of one seed share one stream)."""
    long = native.normals(99, 3_000_001)
    for n in (1, 131_071, 1_048_577, 2_999_999):
        assert np.array_equal(native.normals(99, n), long[:n])
