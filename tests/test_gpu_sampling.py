"""GPU on-device sampler against its CPU restatement (oracle): uniform and
constant columns bit for bit (the same fused multiply-add), log-uniform and
normal columns to a few ulps (log/exp/sqrt/cos of the GPU math library
against the host's)."""
import numpy as np
import pytest
import torch

from sodecl_amd.sampling import sample, table
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols,seed", [(1000, 70, 3), (37, 5, 0), (65536, 8, 2 ** 63 + 5)])
def test_sampler_matches_oracle(rows, cols, seed):
    kinds = ["uniform", "log_uniform", "normal", "constant"]
    rng = np.random.default_rng(cols)
    dists = []
    for c in range(cols):
        k = kinds[c % 4]
        if k == "uniform":
            dists.append((k, rng.uniform(-2, 0), rng.uniform(0.5, 3)))
        elif k == "log_uniform":
            lo = rng.uniform(0.1, 1)
            dists.append((k, lo, lo * rng.uniform(1.5, 4)))
        elif k == "normal":
            dists.append((k, rng.normal(), rng.uniform(0.01, 2)))
        else:
            dists.append((k, rng.normal()))
    dev = sample(dists, rows, seed).cpu().numpy()
    ref = Oracle("oculomotor").sample(table(dists), rows, seed)
    for c, d in enumerate(dists):
        if d[0] in ("uniform", "constant"):
            assert np.array_equal(dev[:, c], ref[:, c]), c
        else:
            scale = max(abs(d[1]), abs(d[2]))
            np.testing.assert_allclose(dev[:, c], ref[:, c], rtol=0, atol=8e-16 * scale * 8)


def test_sampler_writes_into_given_tensor_on_the_current_stream():
    out = torch.full((100, 3), -1.0, dtype=torch.float64, device="cuda:0")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        r = sample([("uniform", 0, 1), ("constant", 7.0), ("normal", 0, 1)], 100, 11, out=out)
    s.synchronize()
    assert r.data_ptr() == out.data_ptr()
    o = out.cpu().numpy()
    assert (o[:, 0] > 0).all() and (o[:, 0] < 1).all() and (o[:, 1] == 7.0).all()
