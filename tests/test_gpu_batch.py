"""Batched implicit integration (one persistent launch over orbits of
several sizes and step counts) against the CPU oracle run orbit by orbit.

The batch solves each Newton iteration's tridiagonal system by partitions
and cyclic reduction instead of the reference's single sweep, which changes
rounding only: snapshots agree with the oracle to 1e-10 relative (per orbit
and equation scale) and Newton update counts agree exactly except where a
residual norm lands within rounding of the stopping threshold.
"""
import numpy as np
import pytest
import torch

from sodecl_amd import configs, sodecl_interface as si
from oracle.oracle import Oracle
from tests.parity import assert_traj_close

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def run_batch(cfg):
    b = si.Batch(cfg["system"], cfg["solver"], cfg["dt"], cfg["ksteps"], cfg["npar"], cfg["class_neq"], 0,
                 {"newton_max_iter": cfg["newton_max_iter"]})
    b.setup(cfg["cls"], cfg["steps"])
    yoff, ooff = b.offsets()
    y0 = torch.tensor(cfg["y0"], device=DEV)
    p = torch.tensor(cfg["params"], device=DEV)
    out = torch.full((b.output_size(),), float("nan"), dtype=torch.float64, device=DEV)
    nu = torch.zeros(cfg["orbits"], dtype=torch.int64, device=DEV)
    b.run_device(y0.data_ptr(), p.data_ptr(), out.data_ptr(), nu.data_ptr(), 0)
    torch.cuda.synchronize()
    return out.cpu().numpy(), nu.cpu().numpy(), yoff, ooff


def check_against_oracle(cfg, out, nu, yoff, ooff, orbits):
    o = Oracle("brusselator")
    solver = {3: "ie", 4: "im"}[cfg["solver"]]
    mism = 0
    for k in orbits:
        n = cfg["class_neq"][cfg["cls"][k]]
        y0 = cfg["y0"][yoff[k]:yoff[k] + n][None]
        nl = int(cfg["steps"][k]) // cfg["ksteps"]
        ref = o.run(solver, y0, cfg["params"][k][None], cfg["dt"], 0.0, cfg["ksteps"], nlaunch=nl,
                    newton_maxit=cfg["newton_max_iter"], use_band=True)
        got = out[ooff[k]:ooff[k] + (nl + 1) * n].reshape(nl + 1, n, 1)
        assert_traj_close(got, ref["out"], what=f"orbit {k} (neq {n}, {cfg['steps'][k]} steps)")
        mism += int(ref["newton_updates"][0] != nu[k])
    return mism


@pytest.mark.parametrize("solver", [3, 4])
def test_batch_mixed_sizes_vs_oracle(solver):
    cfg = configs.brusselator_batch(orbits=300, seed=2, solver=solver, smin=100, smax=400, ksteps=50)
    out, nu, yoff, ooff = run_batch(cfg)
    assert np.all(np.isfinite(out))
    mism = check_against_oracle(cfg, out, nu, yoff, ooff, range(0, 300, 7))
    assert mism <= 2


@pytest.mark.parametrize("cells", [(3,), (5, 40), (100,), (33, 64)])
def test_batch_odd_sizes(cells):
    """Partition edge cases: fewer rows than lanes, partitions of one row,
    sizes not divisible by the wavefront width, a single class."""
    cfg = configs.brusselator_batch(orbits=40, seed=3, solver=3, cells=cells, smin=50, smax=200, ksteps=50,
                                    newton_max_iter=1000)
    out, nu, yoff, ooff = run_batch(cfg)
    check_against_oracle(cfg, out, nu, yoff, ooff, range(40))


@pytest.mark.parametrize("cells", [(16,), (16, 256), (8, 64, 256)])
def test_batch_items_with_unequal_lengths(cells):
    """Orbits that share a work item (up to four small orbits run in lockstep
    on one wavefront) but need very different step counts, zero included,
    with orbit counts that leave items partly empty; the second run of the
    same batch (work queue reset, no state kept) is bit-identical."""
    orbits = 11
    cfg = configs.brusselator_batch(orbits=orbits, seed=6, solver=4, cells=cells, smin=50, smax=100, ksteps=50,
                                    newton_max_iter=1000)
    cfg["steps"] = np.array([0, 1000, 50, 0, 600, 50, 950, 0, 100, 1000, 50], dtype=np.int32)[:orbits]
    out, nu, yoff, ooff = run_batch(cfg)
    zero = [k for k in range(orbits) if cfg["steps"][k] == 0]
    for k in zero:
        n = cfg["class_neq"][cfg["cls"][k]]
        np.testing.assert_array_equal(out[ooff[k]:ooff[k] + n], cfg["y0"][yoff[k]:yoff[k] + n])
        assert nu[k] == 0
    assert check_against_oracle(cfg, out, nu, yoff, ooff, [k for k in range(orbits) if k not in zero]) <= 1
    out2, nu2, _, _ = run_batch(cfg)
    np.testing.assert_array_equal(out2, out)
    np.testing.assert_array_equal(nu2, nu)


def test_batch_reference_newton_cap():
    """With the reference's cap (1000) the Newton iteration converges and the
    update counts match the oracle's."""
    cfg = configs.brusselator_batch(orbits=60, seed=4, solver=3, smin=50, smax=150, ksteps=50, newton_max_iter=1000)
    out, nu, yoff, ooff = run_batch(cfg)
    assert check_against_oracle(cfg, out, nu, yoff, ooff, range(60)) <= 1


def test_batch_rejects_bad_input():
    cfg = configs.brusselator_batch(orbits=10, seed=5)
    b = si.Batch(cfg["system"], 3, cfg["dt"], cfg["ksteps"], 3, cfg["class_neq"], 0, {})
    with pytest.raises(RuntimeError, match="multiples"):
        b.setup(cfg["cls"], cfg["steps"] + 1)
    with pytest.raises(RuntimeError, match="class"):
        b.setup(np.full(10, 7, dtype=np.int32), cfg["steps"])
    with pytest.raises(RuntimeError, match="implicit"):
        si.Batch(cfg["system"], 2, cfg["dt"], cfg["ksteps"], 3, cfg["class_neq"], 0, {})
