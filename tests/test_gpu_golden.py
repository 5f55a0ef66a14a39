"""GPU RHS against the committed regression vectors (no oracle at run time)."""
import os

import numpy as np
import pytest

from casegen import f3d, fp64_rhs_check
from golden.make_golden import golden_cases

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rhs_golden.npz")
NG = 4


@pytest.mark.parametrize("name", ["tgv8", "ramp12"])
def test_gpu_rhs_matches_golden(name):
    import torch
    gold = np.load(GOLDEN, allow_pickle=False)
    c = golden_cases()[name]
    p = c.params
    ctx = f3d.Flame3D(p)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    R = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    phi = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    ctx.rhs(R, d(c.Q), [d(m) for m in c.metrics], d(c.J), phi)
    masks = [torch.zeros(p.mask_shape(k), dtype=torch.uint8, device="cuda:0") for k in range(3)]
    ctx.schemeMask(*masks, phi)
    torch.cuda.synchronize()
    s = (slice(NG, -NG),) * 3
    assert np.array_equal(phi.cpu().numpy()[s], gold[f"{name}_phi"])
    for k in range(3):
        assert np.array_equal(masks[k].cpu().numpy(), gold[f"{name}_mask{k}"])
    # the committed vectors pin the oracle; the GPU is held to the long-double bar on the same input
    fp64_rhs_check(f"golden {name}", R.cpu().numpy(), c)
