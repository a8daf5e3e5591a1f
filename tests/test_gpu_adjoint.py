"""GPU discrete adjoint against the FP64 forward-sensitivity reference
(oracle), central finite differences, and its own algebraic properties.

Tolerances: gradients vs forward sensitivity 1e-9 relative to the largest
gradient entry (both are exact derivatives of the same discrete map; they
differ by rounding); vs central differences (h = 1e-6) 1e-6 relative.
"""
import numpy as np
import pytest
import torch

from sodecl_amd import configs
from sodecl_amd.torch_ops import Integrator
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def make(n, orbits, solver, seed, tspan=0.5, ksteps=5, dt=0.01, group=0, ckpt_every=0):
    cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=seed, dt=dt, tspan=tspan, ksteps=ksteps)
    opts = {"schedule": "group"}
    if group:
        opts["group_size"] = group
    integ = Integrator(cfg["system"], solver, orbits, n, n + 1, dt, tspan, ksteps, ckpt_every=ckpt_every, **opts)
    y0 = torch.tensor(cfg["y0"], device=DEV, requires_grad=True)
    p = torch.tensor(cfg["params"], device=DEV, requires_grad=True)
    return cfg, integ, y0, p


def grads(integ, y0, p, cot):
    y0.grad = None
    p.grad = None
    traj = integ(y0, p)
    (traj * cot).sum().backward()
    torch.cuda.synchronize()
    return y0.grad.cpu().numpy().copy(), p.grad.cpu().numpy().copy()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("n,group", [(6, 8), (16, 16), (20, 8), (64, 64), (32, 32), (3, 4), (5, 2), (6, 64), (10, 0)])
def test_adjoint_vs_forward_sensitivity(solver, n, group):
    """Group sizes cover every pairwise layout: lane exchange (N <= G < 2N),
    the LDS matrix with its pairs spread over spare lanes (G >= 2N, and the
    group widened automatically for this small launch when G is 0), and
    the LDS matrix for groups narrower than N."""
    orbits = 37
    cfg, integ, y0, p = make(n, orbits, solver, seed=n, group=group)
    rng = np.random.default_rng(n)
    cot_np = rng.normal(size=(integ.nsnap, n, orbits))
    cot = torch.tensor(cot_np, device=DEV)
    gy, gp = grads(integ, y0, p, cot)
    if group == 0:
        assert integ.session.launch_config()[1] == 64
    k = 5 if n == 64 else orbits
    ry, rp = Oracle("kuramoto").sensitivity_gradient({1: "euler", 2: "rk4"}[solver], cfg["y0"], cfg["params"], cot_np,
                                                     cfg["dt"], cfg["tspan"], cfg["ksteps"], orbit_range=(0, k))
    assert rel(gy[:k], ry) < 1e-9, rel(gy[:k], ry)
    assert rel(gp[:k], rp) < 1e-9, rel(gp[:k], rp)


def test_adjoint_linear_in_cotangent():
    cfg, integ, y0, p = make(12, 50, 2, seed=3, group=16)
    rng = np.random.default_rng(1)
    c1 = torch.tensor(rng.normal(size=(integ.nsnap, 12, 50)), device=DEV)
    c2 = torch.tensor(rng.normal(size=(integ.nsnap, 12, 50)), device=DEV)
    g1 = grads(integ, y0, p, c1)
    g2 = grads(integ, y0, p, c2)
    g3 = grads(integ, y0, p, 2.0 * c1 - 3.0 * c2)
    for a, b, c in zip(g1, g2, g3):
        np.testing.assert_allclose(c, 2.0 * a - 3.0 * b, rtol=0, atol=1e-12 * np.abs(c).max())


def test_checkpoint_interval_does_not_change_gradients():
    out = []
    for every in (1, 7, 13, 50):
        cfg, integ, y0, p = make(10, 40, 2, seed=4, group=16, ckpt_every=every)
        cot = torch.ones((integ.nsnap, 10, 40), dtype=torch.float64, device=DEV)
        out.append(grads(integ, y0, p, cot))
    for g in out[1:]:
        np.testing.assert_array_equal(g[0], out[0][0])
        np.testing.assert_array_equal(g[1], out[0][1])


def test_forward_with_checkpoints_is_forward():
    cfg, integ, y0, p = make(16, 33, 2, seed=6, group=16)
    a = integ(y0.detach(), p.detach())
    b = integ._forward(y0.detach(), p.detach(), checkpoints=True)[0]
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def sincos_loss(traj, target):
    return ((torch.sin(traj) - torch.sin(target)) ** 2 + (torch.cos(traj) - torch.cos(target)) ** 2).mean()


def test_adjoint_vs_central_differences_sincos_loss():
    """The fitting loss of the benchmark (sin/cos misfit against a target
    trajectory from ground-truth parameters), N=64, RK4, checked by central
    differences with h = 1e-6 on K, one w_i and one theta0_i of two orbits."""
    n, orbits = 64, 4
    cfg, integ, y0, p = make(n, orbits, 2, seed=11, tspan=2.0, ksteps=10, group=64)
    with torch.no_grad():
        truth = p.detach().clone()
        truth[:, 0] *= 1.3
        truth[:, 1:] += 0.05
        target = integ(y0.detach(), truth)
    loss = sincos_loss(integ(y0, p), target)
    loss.backward()
    torch.cuda.synchronize()
    gy = y0.grad.cpu().numpy()
    gp = p.grad.cpu().numpy()
    h = 1e-6

    def L(yv, pv):
        with torch.no_grad():
            return float(sincos_loss(integ(yv, pv), target))

    for o in (0, 3):
        for kind, idx in (("p", 0), ("p", 17), ("y", 40)):
            yv = y0.detach().clone()
            pv = p.detach().clone()
            tgt = pv if kind == "p" else yv
            tgt[o, idx] += h
            lp = L(yv, pv)
            tgt[o, idx] -= 2 * h
            lm = L(yv, pv)
            fd = (lp - lm) / (2 * h)
            an = (gp if kind == "p" else gy)[o, idx]
            assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-3 * float(loss)), (kind, o, idx, an, fd)


def make_implicit(system, solver, orbits, seed, tol=1e-14, ckpt_every=0):
    rng = np.random.default_rng(seed)
    if system == "oculomotor":
        n, npar = 6, 6
        cfg = configs.oculomotor_rk4(orbits=orbits, seed=seed, dt=1e-4, tspan=0.02, ksteps=20)
        y0 = cfg["y0"] + rng.normal(0, 0.05, (orbits, n))
        y0[:, 5] += 1.0
        p = cfg["params"]
    else:
        n, npar = 5, 6
        cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=seed, dt=0.01, tspan=1.0, ksteps=10)
        y0, p = cfg["y0"], cfg["params"]
    integ = Integrator(cfg["system"], solver, orbits, n, npar, cfg["dt"], cfg["tspan"], cfg["ksteps"],
                       ckpt_every=ckpt_every, newton_tol=tol)
    return cfg, integ, y0, p


@pytest.mark.parametrize("system", ["oculomotor", "kuramoto"])
@pytest.mark.parametrize("solver", [3, 4])
def test_implicit_adjoint_vs_forward_sensitivity(system, solver):
    """Implicit Euler / midpoint adjoint (transposed Newton system at the
    converged solution, tridiagonal sweeps) against the oracle's forward
    sensitivity (dense solve of the tangent system), both with the Newton
    threshold at 1e-14 so that the two integrations converge to the same
    states; 1e-9 relative to the largest gradient entry."""
    orbits = 29
    cfg, integ, y0, p = make_implicit(system, solver, orbits, seed=solver)
    yt = torch.tensor(y0, device=DEV, requires_grad=True)
    pt = torch.tensor(p, device=DEV, requires_grad=True)
    rng = np.random.default_rng(solver + 10)
    cot_np = rng.normal(size=(integ.nsnap, integ.neq, orbits))
    gy, gp = grads(integ, yt, pt, torch.tensor(cot_np, device=DEV))
    name = {3: "ie", 4: "im"}[solver]
    ry, rp = Oracle(system).sensitivity_gradient(name, y0, p, cot_np, cfg["dt"], cfg["tspan"], cfg["ksteps"],
                                                 newton_tol=1e-14)
    assert rel(gy, ry) < 1e-9, rel(gy, ry)
    assert rel(gp, rp) < 1e-9, rel(gp, rp)


@pytest.mark.parametrize("system,solver", [("oculomotor", 3), ("kuramoto", 4)])
def test_implicit_adjoint_vs_central_differences(system, solver):
    orbits = 7
    cfg, integ, y0, p = make_implicit(system, solver, orbits, seed=2)
    rng = np.random.default_rng(7)
    cot_np = rng.normal(size=(integ.nsnap, integ.neq, orbits))
    cot = torch.tensor(cot_np, device=DEV)
    yt = torch.tensor(y0, device=DEV, requires_grad=True)
    pt = torch.tensor(p, device=DEV, requires_grad=True)
    gy, gp = grads(integ, yt, pt, cot)

    def L(yv, pv):
        with torch.no_grad():
            traj = integ(torch.tensor(yv, device=DEV), torch.tensor(pv, device=DEV))
            return (traj * cot).sum(dim=(0, 1)).cpu().numpy()

    for j in range(integ.neq):
        h = 1e-6
        yp, ym = y0.copy(), y0.copy()
        yp[:, j] += h
        ym[:, j] -= h
        fd = (L(yp, p) - L(ym, p)) / (2 * h)
        assert np.max(np.abs(gy[:, j] - fd)) <= 1e-6 * max(np.abs(gy).max(), 1e-300), (j, gy[:, j], fd)
    for k in range(integ.npar):
        h = 1e-6 * np.abs(p[:, k]).max()
        pp, pm = p.copy(), p.copy()
        pp[:, k] += h
        pm[:, k] -= h
        fd = (L(y0, pp) - L(y0, pm)) / (2 * h)
        assert np.max(np.abs(gp[:, k] - fd)) <= 1e-6 * max(np.abs(gp).max(), 1e-300), (k, gp[:, k], fd)


def test_implicit_adjoint_checkpoint_interval_and_default_threshold():
    """Gradients are bit-identical for every checkpoint interval (each
    segment is recomputed with the forward's own Newton step), also at the
    reference's Newton threshold 1e-07."""
    orbits = 33
    res = []
    for every in (0, 1, 7, 1000):
        cfg, integ, y0, p = make_implicit("kuramoto", 3, orbits, seed=4, tol=0.0, ckpt_every=every)
        yt = torch.tensor(y0, device=DEV, requires_grad=True)
        pt = torch.tensor(p, device=DEV, requires_grad=True)
        cot = torch.ones((integ.nsnap, integ.neq, orbits), dtype=torch.float64, device=DEV)
        res.append(grads(integ, yt, pt, cot))
    for gy, gp in res[1:]:
        assert np.array_equal(gy, res[0][0]) and np.array_equal(gp, res[0][1])


@pytest.mark.parametrize("solver,n,group", [(1, 16, 16), (2, 64, 64), (2, 20, 8), (2, 10, 64)])
def test_stored_stage_adjoint_bit_identical_to_checkpointed(solver, n, group):
    """The adjoint without recomputation (the forward pass stores every
    step's stage vectors) gives the checkpointed adjoint's gradients and the
    same trajectory, bit for bit."""
    orbits = 37
    cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=n + solver, dt=0.01, tspan=0.5, ksteps=5)
    rng = np.random.default_rng(n)
    cot = torch.tensor(rng.normal(size=(11, n, orbits)), device=DEV)
    res = []
    for stored in (False, True):
        integ = Integrator(cfg["system"], solver, orbits, n, n + 1, cfg["dt"], cfg["tspan"], cfg["ksteps"],
                           store_stages=stored, schedule="group", group_size=group)
        y0 = torch.tensor(cfg["y0"], device=DEV, requires_grad=True)
        p = torch.tensor(cfg["params"], device=DEV, requires_grad=True)
        traj = integ(y0, p)
        (traj * cot).sum().backward()
        torch.cuda.synchronize()
        res.append((traj.detach().cpu().numpy(), y0.grad.cpu().numpy(), p.grad.cpu().numpy()))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_stored_stage_adjoint_rejects_a_stale_forward():
    cfg = configs.kuramoto_rk4(orbits=8, n=6, seed=1, dt=0.01, tspan=0.1, ksteps=5)
    integ = Integrator(cfg["system"], 2, 8, 6, 7, cfg["dt"], cfg["tspan"], cfg["ksteps"], store_stages=True,
                       schedule="group")
    y0 = torch.tensor(cfg["y0"], device=DEV, requires_grad=True)
    p = torch.tensor(cfg["params"], device=DEV, requires_grad=True)
    first = integ(y0, p)
    second = integ(y0, p)
    second.sum().backward()
    with pytest.raises(RuntimeError, match="later forward pass"):
        first.sum().backward()
