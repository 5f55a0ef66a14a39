"""Writes tests/golden/rhs_golden.npz: regression vectors of the CPU oracle.

The reference (Julia + CUDA.jl) has no tests, fixtures or golden vectors and cannot run here, so
these vectors are produced by the oracle itself from the seeded case generators and pin it against
unintended change (and the case generators against drift, via an input digest).  They are not an
independent check of correctness; the closed-form tests in test_oracle_invariants.py are.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from casegen import ramp_case, tgv_case  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

NG = 4


def golden_cases():
    return {
        "tgv8": tgv_case(8, spikes=[(2, 3, 4, 20.0), (6, 5, 1, 3.0)], hybrid_phi3=2.5),
        "ramp12": ramp_case(12, 10, 6, jump=0.8, spikes=[(4, 3, 2, 20.0)], hybrid_phi3=2.5),
    }


def digest(case):
    h = hashlib.sha256()
    for a in [case.Q, case.J] + list(case.metrics):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def compute(case):
    p, o = case.params, Oracle(case.params)
    R = np.zeros(p.cell_shape(5))
    phi = np.zeros(p.cell_shape())
    masks = [np.zeros(p.mask_shape(d), np.uint8) for d in range(3)]
    o.rhs(R, case.Q, case.metrics, case.J, phi, masks)
    s = (slice(NG, -NG),) * 3
    return R[(slice(None),) + s].copy(), phi[s].copy(), masks


def main():
    out = {}
    for name, c in golden_cases().items():
        R, phi, masks = compute(c)
        out[f"{name}_R"] = R
        out[f"{name}_phi"] = phi
        for d in range(3):
            out[f"{name}_mask{d}"] = masks[d]
        out[f"{name}_digest"] = np.frombuffer(digest(c).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "rhs_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "rhs_golden.npz"))


if __name__ == "__main__":
    main()
