"""sodecl_adjoint_device_noalloc (the adjoint whose workspace is the
library's static scratch array) against sodecl_adjoint_device (workspace
passed in) on the same checkpoints and cotangent, through the C ABI: two
calls in a row, the second with a smaller input, gradients bit-identical;
the size limit is a clear error, not a launch."""
import ctypes

import numpy as np
import pytest
import torch

import sodecl_amd
from sodecl_amd import configs, sodecl_interface as si

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _lib():
    lib = ctypes.CDLL(sodecl_amd.library_path())
    f = lib.sodecl_adjoint_device_noalloc
    f.argtypes = [ctypes.c_void_p] * 7 + [ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    lib.sodecl_last_error.argtypes = [ctypes.c_void_p]
    lib.sodecl_last_error.restype = ctypes.c_char_p
    return lib


def setup_case(orbits, n, solver, seed, tspan=0.6, ksteps=5):
    cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=seed, dt=0.01, tspan=tspan, ksteps=ksteps)
    s = si.Session(cfg["system"], solver, orbits, n, n + 1, 0, cfg["dt"], tspan, ksteps, 0, {"schedule": "group"})
    s.setup(cfg["y0"], cfg["params"])
    every = s.default_checkpoint_interval()
    y = torch.tensor(cfg["y0"], device=DEV)
    p = torch.tensor(cfg["params"], device=DEV)
    t = torch.zeros(orbits, dtype=torch.float64, device=DEV)
    traj = torch.empty((s.nsnap, n, orbits), dtype=torch.float64, device=DEV)
    nck = s.checkpoint_count(every)
    ckpt_t = torch.empty((nck, orbits), dtype=torch.float64, device=DEV)
    ckpt_y = torch.empty((nck, n, orbits), dtype=torch.float64, device=DEV)
    s.run_device_checkpointed(t.data_ptr(), y.data_ptr(), p.data_ptr(), 0, traj.data_ptr(), ckpt_t.data_ptr(),
                              ckpt_y.data_ptr(), every, 0)
    rng = np.random.default_rng(seed)
    cot = torch.tensor(rng.normal(size=(s.nsnap, n, orbits)), device=DEV)
    torch.cuda.synchronize()
    return s, every, p, ckpt_t, ckpt_y, cot


def grads_with_workspace(s, every, p, ckpt_t, ckpt_y, cot):
    orbits, n = cot.shape[2], cot.shape[1]
    gy = torch.full((orbits, n), float("nan"), dtype=torch.float64, device=DEV)
    gp = torch.full((orbits, n + 1), float("nan"), dtype=torch.float64, device=DEV)
    work = torch.empty(s.adjoint_workspace_size(every), dtype=torch.float64, device=DEV)
    s.adjoint_device(ckpt_t.data_ptr(), ckpt_y.data_ptr(), p.data_ptr(), cot.data_ptr(), gy.data_ptr(),
                     gp.data_ptr(), work.data_ptr(), every, 0)
    torch.cuda.synchronize()
    return gy.cpu(), gp.cpu()


def grads_noalloc(lib, s, every, p, ckpt_t, ckpt_y, cot):
    orbits, n = cot.shape[2], cot.shape[1]
    gy = torch.full((orbits, n), float("nan"), dtype=torch.float64, device=DEV)
    gp = torch.full((orbits, n + 1), float("nan"), dtype=torch.float64, device=DEV)
    ok = lib.sodecl_adjoint_device_noalloc(s.handle(), ckpt_t.data_ptr(), ckpt_y.data_ptr(), p.data_ptr(),
                                           cot.data_ptr(), gy.data_ptr(), gp.data_ptr(), every, None)
    assert ok == 1, lib.sodecl_last_error(s.handle())
    torch.cuda.synchronize()
    return gy.cpu(), gp.cpu()


@pytest.mark.parametrize("solver", [2, 1])
def test_noalloc_adjoint_matches_workspace_adjoint(solver):
    lib = _lib()
    # first a larger case, then a smaller one on the same GPU: the scratch
    # keeps the first call's rows, which the second must not read
    for orbits, n, seed in ((300, 16, 1), (37, 10, 2)):
        case = setup_case(orbits, n, solver, seed)
        s = case[0]
        assert s.adjoint_workspace_size(case[1]) <= 8 << 20
        want = grads_with_workspace(*case)
        got = grads_noalloc(lib, *case)
        assert torch.isfinite(want[0]).all() and torch.isfinite(want[1]).all()
        assert torch.equal(got[0], want[0]), (orbits, n)
        assert torch.equal(got[1], want[1]), (orbits, n)


def test_noalloc_adjoint_rejects_oversized_workspace():
    lib = _lib()
    s, every, p, ckpt_t, ckpt_y, cot = setup_case(1024, 64, 2, 3, tspan=20.0, ksteps=10)
    assert s.adjoint_workspace_size(every) > 8 << 20
    gy = torch.empty((1024, 64), dtype=torch.float64, device=DEV)
    gp = torch.empty((1024, 65), dtype=torch.float64, device=DEV)
    ok = lib.sodecl_adjoint_device_noalloc(s.handle(), ckpt_t.data_ptr(), ckpt_y.data_ptr(), p.data_ptr(),
                                           cot.data_ptr(), gy.data_ptr(), gp.data_ptr(), every, None)
    assert ok == 0
    assert b"exceeds" in lib.sodecl_last_error(s.handle())
