"""Edge cases of the launch sequencing on the GPU, each against the oracle
run the same way: no steps at all (tspan = 0: only the initial snapshot),
a single orbit, orbit counts that leave partial blocks and partial groups,
one step per segment, a single segment longer than tspan (the launch-count
rule rounds up), and the None output type (state advanced, nothing
returned)."""
import numpy as np
import pytest

from sodecl_amd import configs, sodecl_interface as si
from oracle.oracle import Oracle
from tests.parity import assert_traj_close

pytestmark = pytest.mark.gpu


def ocu(orbits, dt=2e-6, tspan=0.002, ksteps=100, seed=1, solver=2):
    cfg = configs.oculomotor_rk4(orbits=orbits, seed=seed, dt=dt, tspan=tspan, ksteps=ksteps)
    cfg["solver"] = solver
    cfg["y0"][:, 5] += 1.0
    return cfg


def run(cfg, **opts):
    s = configs.session_for(cfg, **opts)
    s.setup(cfg["y0"], cfg["params"], None, cfg.get("counters"))
    return s, s.run().reshape(s.nsnap, cfg["neq"], cfg["orbits"])


def oracle(cfg, system="oculomotor"):
    name = {0: "em", 1: "euler", 2: "rk4", 3: "ie", 4: "im"}[cfg["solver"]]
    return Oracle(system).run(name, cfg["y0"], cfg["params"], cfg["dt"], cfg["tspan"], cfg["ksteps"],
                              counters=cfg.get("counters"), nnoi=cfg["nnoi"])["out"]


@pytest.mark.parametrize("schedule", ["lane", "team", "plain"])
def test_zero_time_span_returns_the_initial_state(schedule):
    cfg = ocu(77, tspan=0.0)
    s, out = run(cfg, schedule=schedule)
    assert s.nsnap == 1 and s.launches == 0
    assert np.array_equal(out[0], cfg["y0"].T)


@pytest.mark.parametrize("orbits", [1, 63, 65, 129, 1000])
@pytest.mark.parametrize("schedule", ["lane", "team"])
def test_partial_blocks(orbits, schedule):
    cfg = ocu(orbits, seed=orbits)
    _, out = run(cfg, schedule=schedule)
    assert_traj_close(out, oracle(cfg), what=f"{schedule} orbits={orbits}")


@pytest.mark.parametrize("orbits,group,block", [(1, 16, 256), (5, 16, 64), (33, 8, 256)])
def test_partial_groups(orbits, group, block):
    cfg = configs.kuramoto_rk4(orbits=orbits, n=10, seed=orbits, dt=0.01, tspan=0.5, ksteps=5)
    _, out = run(cfg, schedule="group", group_size=group, block_size=block)
    ref = Oracle("kuramoto").run("rk4", cfg["y0"], cfg["params"], cfg["dt"], cfg["tspan"], cfg["ksteps"])["out"]
    assert_traj_close(out, ref, what=f"group orbits={orbits} G={group} B={block}")


@pytest.mark.parametrize("ksteps,tspan", [(1, 2e-5), (7, 0.0001), (1000, 0.0003)])
def test_segment_lengths_and_launch_rounding(ksteps, tspan):
    """One step per segment; a segment length that does not divide the span;
    one segment longer than the whole span (one launch of ksteps steps)."""
    cfg = ocu(100, tspan=tspan, ksteps=ksteps)
    s, out = run(cfg)
    ref = oracle(cfg)
    assert out.shape == ref.shape == (si.num_launches(tspan, 2e-6, ksteps) + 1, 6, 100)
    assert_traj_close(out, ref, what=f"ksteps={ksteps}")


def test_output_none_integrates_without_output():
    """Output type None (sodecl_create with output_type 0): nothing is
    returned, but the run happens: the Euler-Maruyama counters advance by
    one Threefry call per step (two noise terms) on every run."""
    import ctypes
    import sodecl_amd
    cfg = ocu(40, solver=0)
    p = np.ascontiguousarray(np.concatenate([cfg["params"], np.full((40, 2), 0.1)], 1))
    y0 = np.ascontiguousarray(cfg["y0"])
    lib = ctypes.CDLL(sodecl_amd.library_path())
    lib.sodecl_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6
    for name in ("sodecl_set_y0", "sodecl_set_params", "sodecl_set_rng_counters", "sodecl_get_rng_counters"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    for name in ("sodecl_setup", "sodecl_run", "sodecl_destroy", "sodecl_output_size"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.sodecl_output_size.restype = ctypes.c_size_t
    ctx = ctypes.c_void_p()
    assert lib.sodecl_create(ctypes.byref(ctx), b"kernels", cfg["system"].encode(), 0, cfg["dt"], cfg["tspan"],
                             cfg["ksteps"], 6, 8, 2, 40, 0) == 1
    c0 = np.arange(40, dtype=np.float64) * 1e6
    lib.sodecl_set_y0(ctx, y0.ctypes.data)
    lib.sodecl_set_params(ctx, p.ctypes.data)
    assert lib.sodecl_set_rng_counters(ctx, c0.ctypes.data) == 1
    assert lib.sodecl_setup(ctx) == 1
    steps = si.num_launches(cfg["tspan"], cfg["dt"], cfg["ksteps"]) * cfg["ksteps"]
    for k in (1, 2):
        assert lib.sodecl_run(ctx) == 1
        assert lib.sodecl_output_size(ctx) == 0
        c = np.empty(40)
        assert lib.sodecl_get_rng_counters(ctx, c.ctypes.data) == 1
        assert np.array_equal(c, c0 + k * steps)
    lib.sodecl_destroy(ctx)
