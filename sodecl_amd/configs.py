"""Seeded synthetic workloads of the benchmark configurations.

Each function returns a dict with the integration settings and numpy inputs
in the reference's host layouts (y0[orbits][neq], params[orbits][npar]).
All draws use numpy.random.default_rng(seed) (PCG64).

  kuramoto_euler      Kuramoto N=10, explicit Euler, dt=0.01, 1000 steps,
                      snapshot every 10 steps, 1024 orbits; K~U(0,5),
                      w_i~U(0.5,1.5), theta0_i~U(0,2pi)
  oculomotor_rk4      6-state oculomotor model, classic RK4, dt=1e-3, 20000
                      steps, snapshot every 100, 65536 orbits; each parameter
                      log-uniform within x0.5..x2 of (120, 1.5, 0.0045, 0.05,
                      600, 9); initial state zeros plus N(0, 1e-3) noise
  oculomotor_rk4_plain  the same with a plain-style plugin (a user's source)
  kuramoto_em         Kuramoto N=32, Euler-Maruyama, additive noise
                      sigma~U(0,0.5) per orbit, dt=1e-3, 10000 steps, snapshot
                      every 100, 262144 orbits; K~U(0,5), w_i~N(1,0.1);
                      Threefry counters: orbit id in the first counter word,
                      seed*2^32 + (step*16 + call) in the second
  brusselator_batch   Brusselator chains of 16/64/256 cells (1:1:1), implicit
                      Euler or midpoint, 32768 orbits, 500..5000 steps each
  kuramoto_adjoint    Kuramoto N=64, RK4, dt=0.01, 2000 steps, 16384 orbits,
                      target trajectory from seeded ground truth
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYSTEMS = os.path.join(HERE, "systems")

OCULOMOTOR_DEFAULTS = np.array([120.0, 1.5, 0.0045, 0.05, 600.0, 9.0])


def kuramoto_euler(orbits=1024, n=10, seed=0):
    rng = np.random.default_rng(seed)
    K = rng.uniform(0, 5, orbits)
    w = rng.uniform(0.5, 1.5, (orbits, n))
    theta0 = rng.uniform(0, 2 * np.pi, (orbits, n))
    return dict(name="kuramoto_euler", system=os.path.join(SYSTEMS, "kuramoto.cl"), solver=1, neq=n,
                npar=n + 1, nnoi=0, orbits=orbits, dt=0.01, tspan=10.0, ksteps=10,
                y0=theta0, params=np.concatenate([K[:, None], w], 1), counters=None,
                flops_per_orbit_step=2 * n * n + 3 * n + 2 * n + 1 + 20 * n * n)


def oculomotor_rk4(orbits=65536, seed=0, dt=1e-3, tspan=20.0, ksteps=100):
    rng = np.random.default_rng(seed)
    scale = np.exp(rng.uniform(np.log(0.5), np.log(2.0), (orbits, 6)))
    params = OCULOMOTOR_DEFAULTS[None, :] * scale
    y0 = rng.normal(0.0, 1e-3, (orbits, 6))
    return dict(name="oculomotor_rk4", system=os.path.join(SYSTEMS, "oculomotor.cl"), solver=2, neq=6,
                npar=6, nnoi=0, orbits=orbits, dt=dt, tspan=tspan, ksteps=ksteps, y0=y0, params=params,
                counters=None, flops_per_orbit_step=384)


def oculomotor_rk4_plain(orbits=65536, seed=0, dt=1e-3, tspan=20.0, ksteps=100):
    """oculomotor_rk4 with the plain-style plugin (systems/oculomotor_plain.cl:
    branches per burst-curve side, IEEE divisions), as a user would write it."""
    cfg = oculomotor_rk4(orbits=orbits, seed=seed, dt=dt, tspan=tspan, ksteps=ksteps)
    cfg["name"] = "oculomotor_rk4_plain_plugin"
    cfg["system"] = os.path.join(SYSTEMS, "oculomotor_plain.cl")
    return cfg


def kuramoto_em(orbits=262144, n=32, seed=0, tspan=10.0, ksteps=100):
    rng = np.random.default_rng(seed)
    K = rng.uniform(0, 5, orbits)
    w = rng.normal(1.0, 0.1, (orbits, n))
    theta0 = rng.uniform(0, 2 * np.pi, (orbits, n))
    sigma = rng.uniform(0, 0.5, orbits)
    params = np.concatenate([K[:, None], w, np.repeat(sigma[:, None], n, 1)], 1)
    counters = np.full(orbits, float(seed) * 2.0 ** 32)
    calls = (n + 1) // 2
    # RHS 2N^2+3N, noise calls*(4 mul + sqrt) + N, stoch N, update 2N, t 1;
    # transcendentals: N^2 sin + calls*(log + sincos)
    flops = (2 * n * n + 3 * n) + calls * 9 + n + 2 * n + 2 * n + 1 + 20 * (n * n + 3 * calls)
    return dict(name="kuramoto_em", system=os.path.join(SYSTEMS, "kuramoto.cl"), solver=0, neq=n,
                npar=2 * n + 1, nnoi=n, orbits=orbits, dt=1e-3, tspan=tspan, ksteps=ksteps, y0=theta0,
                params=params, counters=counters, flops_per_orbit_step=flops)


def kuramoto_rk4(orbits=16384, n=64, seed=0, dt=0.01, tspan=20.0, ksteps=10):
    rng = np.random.default_rng(seed)
    K = rng.uniform(0, 5, orbits)
    w = rng.normal(1.0, 0.1, (orbits, n))
    theta0 = rng.uniform(0, 2 * np.pi, (orbits, n))
    rhs = 2 * n * n + 3 * n + 20 * n * n
    return dict(name="kuramoto_rk4", system=os.path.join(SYSTEMS, "kuramoto.cl"), solver=2, neq=n,
                npar=n + 1, nnoi=0, orbits=orbits, dt=dt, tspan=tspan, ksteps=ksteps, y0=theta0,
                params=np.concatenate([K[:, None], w], 1), counters=None,
                flops_per_orbit_step=4 * rhs + 14 * n + 4)


def brusselator_initial(cells, B, A=1.0):
    x = (np.arange(cells) + 1.0) / (cells + 1.0)
    y = np.empty(2 * cells)
    y[0::2] = A + 0.5 * np.sin(2 * np.pi * x)
    y[1::2] = B / A
    return y


def session_for(cfg, output_type=1, **options):
    """A Session of the workload; output_type 1 = the trajectory returned to
    host memory (ARRAY), 0 = none."""
    from . import sodecl_interface
    return sodecl_interface.Session(cfg["system"], cfg["solver"], cfg["orbits"], cfg["neq"], cfg["npar"],
                                    cfg["nnoi"], cfg["dt"], cfg["tspan"], cfg["ksteps"], 0, options,
                                    output_type=output_type)


def warm_jit_cache(arch="gfx950"):
    """Compiles the programs of the bundled workloads into the on-disk cache."""
    from . import sodecl_interface
    specs = [
        # system, solver, orbits, neq, npar, nnoi, dt, tspan, ksteps, options
        ("kuramoto.cl", 1, 1024, 10, 11, 0, 0.01, 10.0, 10, {}),
        ("oculomotor.cl", 2, 65536, 6, 6, 0, 1e-3, 20.0, 100, {}),
        ("oculomotor_plain.cl", 2, 65536, 6, 6, 0, 1e-3, 20.0, 100, {}),
        ("kuramoto.cl", 0, 262144, 32, 65, 32, 1e-3, 10.0, 100, {}),
        ("kuramoto.cl", 2, 16384, 64, 65, 0, 0.01, 20.0, 10, {}),
    ]
    for name, solver, orbits, neq, npar, nnoi, dt, tspan, ksteps, opts in specs:
        s = sodecl_interface.Session(os.path.join(SYSTEMS, name), solver, orbits, neq, npar, nnoi, dt, tspan,
                                     ksteps, 0, opts)
        s.compile_only(arch)


def brusselator_batch(orbits=32768, seed=0, solver=3, cells=(16, 64, 256), dt=0.01, ksteps=50, smin=500,
                      smax=5000, newton_max_iter=5):
    """Brusselator chains of 16/64/256 cells (neq 32/128/512) in equal
    shares, each orbit with its own step count drawn uniformly from the
    multiples of ksteps in [smin, smax]; A = 1, B ~ U(2, 4), a ~ U(0.01, 1).
    Packed state: orbit o's initial state at y0[yoff[o]]."""
    rng = np.random.default_rng(seed)
    cls = rng.permutation(np.arange(orbits) % len(cells)).astype(np.int32)
    steps = (ksteps * rng.integers(smin // ksteps, smax // ksteps + 1, orbits)).astype(np.int32)
    B = rng.uniform(2.0, 4.0, orbits)
    a = rng.uniform(0.01, 1.0, orbits)
    params = np.stack([np.ones(orbits), B, a], 1)
    class_neq = [2 * c for c in cells]
    y0 = np.concatenate([brusselator_initial(cells[c], B[o]) for o, c in enumerate(cls)])
    return dict(name="brusselator_batch", system=os.path.join(SYSTEMS, "brusselator.cl"), solver=solver,
                class_neq=class_neq, cls=cls, steps=steps, params=params, y0=y0, dt=dt, ksteps=ksteps, npar=3,
                orbits=orbits, newton_max_iter=newton_max_iter)


def brusselator_flops(neq, steps, updates, solver):
    """Algorithmic FP64 flops of the batched Brusselator run, the count the
    workload is specified with: per state and Newton update 30 (implicit
    Euler) / 33 (implicit midpoint, which also forms the midpoint), per state
    and step 12.5 / 15.5 for the residual check that ends each step's
    iteration; a division or square root counts 1, no transcendentals."""
    neq = np.asarray(neq, dtype=np.float64)
    if solver == 3:
        return float(np.sum(neq * (30.0 * np.asarray(updates) + 12.5 * np.asarray(steps))))
    return float(np.sum(neq * (33.0 * np.asarray(updates) + 15.5 * np.asarray(steps))))
