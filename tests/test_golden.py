"""The oracle reproduces its committed regression vectors (tests/golden/rhs_golden.npz) and the
case generators still produce the same inputs.  CPU only; the GPU against the same vectors is in
test_gpu_golden."""
import os

import numpy as np
import pytest

from golden.make_golden import compute, digest, golden_cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rhs_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("name", ["tgv8", "ramp12"])
def test_oracle_matches_golden(gold, name):
    c = golden_cases()[name]
    assert digest(c) == bytes(gold[f"{name}_digest"]).decode(), "case generator drifted"
    R, phi, masks = compute(c)
    assert np.array_equal(phi, gold[f"{name}_phi"])
    for d in range(3):
        assert np.array_equal(masks[d], gold[f"{name}_mask{d}"])
    assert np.array_equal(R, gold[f"{name}_R"])
    # every scheme is exercised by the two cases together
    seen = set()
    for nm in ("tgv8", "ramp12"):
        for d in range(3):
            seen |= set(np.unique(gold[f"{nm}_mask{d}"]).tolist())
    assert seen == {0, 1, 2, 3}
