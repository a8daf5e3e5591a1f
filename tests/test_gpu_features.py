"""GPU tests of the launch machinery: the autotuner, hipGraph replay and the
pairwise drift form.  None may change results: the tuned configuration is
checked against the oracle at 1e-10, graph replays against the eager
launches and the pairwise form against the per-component form bit for
bit (LDS layouts, j-order sums) or at the parity tolerance (lane exchange,
offset-order sums)."""
import os

import numpy as np
import pytest
import torch

from sodecl_amd import configs
from oracle.oracle import Oracle
from tests.parity import assert_traj_close

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def test_autotune_choice_is_timed_cached_and_correct(tmp_path, monkeypatch):
    monkeypatch.setenv("SODECL_AMD_CACHE_DIR", str(tmp_path))
    cfg = configs.kuramoto_euler(orbits=777, n=20, seed=9)
    cfg["tspan"] = 1.0
    s = configs.session_for(cfg, autotune=1)
    s.setup(cfg["y0"], cfg["params"])
    report = s.autotune_report()
    assert "group" in report and "lane" in report, report
    out = s.run().reshape(s.nsnap, 20, 777)
    ref = Oracle("kuramoto").run("euler", cfg["y0"], cfg["params"], cfg["dt"], cfg["tspan"], cfg["ksteps"])["out"]
    assert_traj_close(out, ref, what=f"tuned {s.launch_config()}")
    s2 = configs.session_for(cfg, autotune=1)
    s2.setup(cfg["y0"], cfg["params"])
    assert s2.autotune_report() == "cached"
    assert any(f.startswith("tune_") for f in os.listdir(tmp_path))


def test_graph_replay_matches_eager_launches():
    cfg = configs.oculomotor_rk4(orbits=300, seed=3, dt=2e-6, tspan=0.004, ksteps=100)
    nsteps_launch = 1   # one launch per ksteps segment: 20 launches recorded in the graph

    def device_run(session, y_init, p, t, traj, stream):
        y = y_init.clone()
        t.zero_()
        session.run_device(t.data_ptr(), y.data_ptr(), p.data_ptr(), 0, traj.data_ptr(), stream)
        return y

    y_init = torch.tensor(cfg["y0"], device=DEV)
    p = torch.tensor(cfg["params"], device=DEV)
    eager = configs.session_for(cfg, schedule="lane", segments_per_launch=nsteps_launch)
    eager.setup(cfg["y0"], cfg["params"])
    t = torch.zeros(300, dtype=torch.float64, device=DEV)
    traj_e = torch.empty((eager.nsnap, 6, 300), dtype=torch.float64, device=DEV)
    ye = device_run(eager, y_init, p, t, traj_e, 0)
    torch.cuda.synchronize()

    graph = configs.session_for(cfg, schedule="lane", segments_per_launch=nsteps_launch, use_graph=1)
    graph.setup(cfg["y0"], cfg["params"])
    y = torch.empty_like(y_init)
    tg = torch.zeros(300, dtype=torch.float64, device=DEV)
    traj_g = torch.empty_like(traj_e)
    for rep in range(4):
        y.copy_(y_init)                      # eager work between replays
        tg.zero_()
        traj_g.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.run_device(tg.data_ptr(), y.data_ptr(), p.data_ptr(), 0, traj_g.data_ptr(), 0)
        torch.cuda.synchronize()
        assert torch.equal(traj_g, traj_e), f"replay {rep}"
        assert torch.equal(y, ye), f"replay {rep}"
    ref = Oracle("oculomotor").run("rk4", cfg["y0"], cfg["params"], cfg["dt"], cfg["tspan"], cfg["ksteps"])["out"]
    assert_traj_close(traj_g.cpu().numpy(), ref, what="graph")


def test_graph_rerecorded_after_recompiling_option():
    """An option that changes the program text (here library_sin) recompiles
    the context's programs; a graph recorded before it held the old modules'
    kernels, so the next device run with the same buffers records anew and
    runs the new program: bit-identical to a fresh context with that option,
    and within the parity bar of the oracle."""
    cfg = configs.kuramoto_euler(orbits=130, n=10, seed=9)
    cfg["solver"] = 2
    cfg["tspan"] = 0.5
    opts = dict(schedule="group", group_size=16, segments_per_launch=1, use_graph=1)
    y_init = torch.tensor(cfg["y0"], device=DEV)
    p = torch.tensor(cfg["params"], device=DEV)
    t = torch.zeros(130, dtype=torch.float64, device=DEV)
    y = torch.empty_like(y_init)

    def device_run(session, traj):
        y.copy_(y_init)
        t.zero_()
        torch.cuda.synchronize()
        session.run_device(t.data_ptr(), y.data_ptr(), p.data_ptr(), 0, traj.data_ptr(), 0)
        torch.cuda.synchronize()
        return traj.clone()

    s = configs.session_for(cfg, **opts)
    s.setup(cfg["y0"], cfg["params"])
    traj = torch.empty((s.nsnap, 10, 130), dtype=torch.float64, device=DEV)
    first = device_run(s, traj)
    s.set_options(dict(opts, library_sin=1))
    second = device_run(s, traj)
    fresh = configs.session_for(cfg, **dict(opts, library_sin=1))
    fresh.setup(cfg["y0"], cfg["params"])
    want = device_run(fresh, torch.empty_like(traj))
    assert torch.equal(second, want)
    assert not torch.equal(first, second)
    ref = Oracle("kuramoto").run("rk4", cfg["y0"], cfg["params"], cfg["dt"], cfg["tspan"], cfg["ksteps"])["out"]
    assert_traj_close(first.cpu().numpy(), ref, what="before the option change")
    assert_traj_close(second.cpu().numpy(), ref, what="after the option change")


@pytest.mark.parametrize("solver,n,group", [(1, 7, 8), (2, 10, 4), (2, 32, 32), (2, 64, 64), (0, 32, 16), (0, 5, 8),
                                            (1, 10, 64), (2, 7, 32), (0, 6, 32)])
def test_pairwise_drift_bit_identical(solver, n, group):
    """The pairwise form of the Kuramoto drift (each sin(y_j - y_i) evaluated
    once, its negation mirrored), in both LDS layouts, sums the same terms
    in the same order as sode_system_i, so trajectories must agree bit for
    bit -- also with groups wider than the system, whose spare lanes
    evaluate pairs too."""
    orbits = 301
    if solver == 0:
        cfg = configs.kuramoto_em(orbits=orbits, n=n, seed=n, tspan=0.2, ksteps=10)
    else:
        cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=n, dt=0.01, tspan=1.0, ksteps=10)
        cfg["solver"] = solver
    outs = []
    for mode in (-1, 3, 2):   # per-component, full matrix, half storage
        s = configs.session_for(cfg, schedule="group", group_size=group, pairwise=mode)
        s.setup(cfg["y0"], cfg["params"], None, cfg.get("counters"))
        outs.append(s.run().copy())
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])
    if solver != 0:
        ref = Oracle("kuramoto").run({1: "euler", 2: "rk4"}[solver], cfg["y0"], cfg["params"], cfg["dt"],
                                     cfg["tspan"], cfg["ksteps"])["out"]
        assert_traj_close(outs[0].reshape(-1, n, orbits), ref, what=f"pairs N={n}")


@pytest.mark.parametrize("solver,n,group", [(1, 7, 8), (2, 32, 32), (2, 64, 64), (2, 33, 64), (2, 3, 4),
                                            (0, 5, 8), (0, 32, 32), (1, 10, 64), (0, 6, 32)])
def test_pairwise_lane_exchange_within_tolerance(solver, n, group):
    """The lane-exchange pairwise form (one component per lane, no LDS
    matrix) sums each row's terms in offset order rather than j order: the
    drift is the per-component drift up to the rounding of that sum, so
    trajectories agree with the per-component form and with the oracle at
    the FP64 parity tolerance, and EM noise counters exactly -- also in
    groups with lanes to spare past the last row."""
    orbits = 301
    if solver == 0:
        cfg = configs.kuramoto_em(orbits=orbits, n=n, seed=n, tspan=0.2, ksteps=10)
        cfg["counters"] = np.random.default_rng(n).integers(0, 2 ** 40, orbits).astype(np.float64)
    else:
        cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=n, dt=0.01, tspan=1.0, ksteps=10)
        cfg["solver"] = solver
    outs, cnts = [], []
    for mode in (-1, 4):
        s = configs.session_for(cfg, schedule="group", group_size=group, pairwise=mode)
        s.setup(cfg["y0"], cfg["params"], None, cfg.get("counters"))
        assert "#define SODECL_USE_PAIRS " + ("0" if mode < 0 else "3") in s.program_source()
        outs.append(s.run().copy().reshape(-1, n, orbits))
        cnts.append(s.counters().copy() if solver == 0 else None)
    assert_traj_close(outs[1], outs[0], what=f"exchange vs per-component N={n}")
    ref = Oracle("kuramoto").run({0: "em", 1: "euler", 2: "rk4"}[solver], cfg["y0"], cfg["params"], cfg["dt"],
                                 cfg["tspan"], cfg["ksteps"], counters=cfg.get("counters"), nnoi=cfg["nnoi"])
    assert_traj_close(outs[1], ref["out"], what=f"exchange vs oracle N={n}")
    if solver == 0:
        np.testing.assert_array_equal(cnts[1], ref["counters"])


def test_pairwise_lane_exchange_adjoint():
    """Discrete adjoint (RK4, group schedule) with the lane-exchange drift:
    its recomputation repeats the forward pass bit for bit, so gradients
    with stored stage vectors equal the checkpointed ones, and both agree
    with the per-component form's gradients at the FP64 parity tolerance."""
    from sodecl_amd.torch_ops import Integrator
    n, orbits = 32, 70
    cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=5, dt=0.01, tspan=0.5, ksteps=5)
    cot = None
    res = []
    for mode, stored in ((-1, False), (4, False), (4, True)):
        integ = Integrator(cfg["system"], 2, orbits, n, n + 1, cfg["dt"], cfg["tspan"], cfg["ksteps"],
                           schedule="group", group_size=32, pairwise=mode, store_stages=stored)
        y0 = torch.tensor(cfg["y0"], device=DEV, requires_grad=True)
        p = torch.tensor(cfg["params"], device=DEV, requires_grad=True)
        if cot is None:
            cot = torch.tensor(np.random.default_rng(5).normal(size=(integ.nsnap, n, orbits)), device=DEV)
        traj = integ(y0, p)
        (traj * cot).sum().backward()
        torch.cuda.synchronize()
        res.append((traj.detach().cpu().numpy(), y0.grad.cpu().numpy(), p.grad.cpu().numpy()))
    for a, b in zip(res[1], res[2]):
        assert np.array_equal(a, b)
    assert_traj_close(res[1][0], res[0][0], what="trajectory")
    for a, b, what in ((res[1][1], res[0][1], "grad y0"), (res[1][2], res[0][2], "grad params")):
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err <= 1e-10, f"{what}: {err:.2e}"


@pytest.mark.parametrize("solver", [1, 2])
def test_pairwise_adjoint_bit_identical(solver):
    from sodecl_amd.torch_ops import Integrator
    n, orbits = 24, 45
    cfg = configs.kuramoto_rk4(orbits=orbits, n=n, seed=4, dt=0.01, tspan=0.5, ksteps=5)
    rng = np.random.default_rng(4)
    res = []
    for mode in (2, -1, 3):
        integ = Integrator(cfg["system"], solver, orbits, n, n + 1, cfg["dt"], cfg["tspan"], cfg["ksteps"],
                           schedule="group", group_size=8, pairwise=mode)
        y0 = torch.tensor(cfg["y0"], device=DEV, requires_grad=True)
        p = torch.tensor(cfg["params"], device=DEV, requires_grad=True)
        cot = torch.tensor(rng.normal(size=(integ.nsnap, n, orbits)) if not res else res[0][3], device=DEV)
        traj = integ(y0, p)
        (traj * cot).sum().backward()
        torch.cuda.synchronize()
        res.append((traj.detach().cpu().numpy(), y0.grad.cpu().numpy(), p.grad.cpu().numpy(), cot.cpu().numpy()))
    for r in res[1:]:
        for a, b in zip(res[0][:3], r[:3]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("orbits,group,block", [(301, 2, 128), (64, 2, 64), (1000, 4, 256)])
def test_team_schedule_bit_identical_to_lane(solver, orbits, group, block):
    """The team schedule steps every orbit with the lane schedule's
    arithmetic; the oculomotor team entry point only moves one of the two
    burst-curve evaluations to the partner lane.  Trajectories, final states
    and continued runs must agree bit for bit, and with the oracle at 1e-10
    (well-conditioned step size)."""
    cfg = configs.oculomotor_rk4(orbits=orbits, seed=orbits + solver, dt=2e-6, tspan=0.004, ksteps=100)
    cfg["solver"] = solver
    cfg["y0"][:, 5] += 1.0
    outs = []
    for opts in (dict(schedule="lane"), dict(schedule="team", group_size=group, block_size=block)):
        s = configs.session_for(cfg, **opts)
        s.setup(cfg["y0"], cfg["params"])
        first = s.run().copy()
        second = s.run().copy()
        outs.append((first, second, s.launch_config()))
    assert outs[1][2] == (4, group, block)
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    ref = Oracle("oculomotor").run({1: "euler", 2: "rk4"}[solver], cfg["y0"], cfg["params"], cfg["dt"],
                                   cfg["tspan"], cfg["ksteps"])["out"]
    assert_traj_close(outs[1][0].reshape(-1, 6, orbits), ref, what=f"team G={group}")


def test_team_schedule_device_run_with_checkpoints_matches_host_run():
    cfg = configs.oculomotor_rk4(orbits=517, seed=2, dt=2e-6, tspan=0.004, ksteps=100)
    cfg["y0"][:, 5] += 1.0
    s = configs.session_for(cfg, schedule="team")
    s.setup(cfg["y0"], cfg["params"])
    host = s.run().reshape(s.nsnap, 6, 517)
    y = torch.tensor(cfg["y0"], device=DEV)
    p = torch.tensor(cfg["params"], device=DEV)
    t = torch.zeros(517, dtype=torch.float64, device=DEV)
    traj = torch.empty((s.nsnap, 6, 517), dtype=torch.float64, device=DEV)
    every = 7
    k = s.checkpoint_count(every)
    ck_t = torch.empty((k, 517), dtype=torch.float64, device=DEV)
    ck_y = torch.empty((k, 6, 517), dtype=torch.float64, device=DEV)
    s.run_device_checkpointed(t.data_ptr(), y.data_ptr(), p.data_ptr(), 0, traj.data_ptr(), ck_t.data_ptr(),
                              ck_y.data_ptr(), every, 0)
    torch.cuda.synchronize()
    assert s.launch_config()[0] == 4
    tr = traj.cpu().numpy()
    assert np.array_equal(tr, host)
    # checkpoint j is the state before step j * every; snapshots are every 100 steps
    cy = ck_y.cpu().numpy()
    for j in range(k):
        if (j * every) % 100 == 0:
            assert np.array_equal(cy[j], tr[(j * every) // 100])
    assert np.allclose(ck_t.cpu().numpy()[3], 3 * every * cfg["dt"], rtol=1e-12)


@pytest.mark.parametrize("solver", [2, 3, 0])
def test_host_share_run_vs_oracle(solver):
    """sodecl_run with a host share: the last 30% of the orbits integrated on
    the host cores (the product's lane integrator compiled for the host)
    while the GPU integrates the rest.  Every orbit within 1e-10 of the
    oracle, the GPU's orbits bit-identical to a GPU-only run, Threefry
    counters exact, and a second run continues from the split state."""
    from sodecl_amd import sodecl_interface as si
    orbits = 200
    explicit = solver in (0, 2)
    dt, tspan, ks = (2e-6, 0.01, 100) if explicit else (1e-5, 0.02, 50)
    cfg = configs.oculomotor_rk4(orbits=orbits, seed=7, dt=dt, tspan=tspan, ksteps=ks)
    cfg["y0"][:, 5] += 1.0
    npar, nnoi, counters = 6, 0, None
    params = cfg["params"]
    if solver == 0:
        rng = np.random.default_rng(7)
        npar, nnoi = 8, 2
        params = np.concatenate([params, rng.uniform(0, 0.2, (orbits, 2))], 1)
        counters = rng.integers(0, 10 ** 11, orbits).astype(np.float64)

    def session(**opts):
        s = si.Session(cfg["system"], solver, orbits, 6, npar, nnoi, dt, tspan, ks, 0, dict(schedule="lane", **opts))
        s.setup(cfg["y0"], params, None, counters)
        return s

    gpu = session()
    a = gpu.run().reshape(gpu.nsnap, 6, orbits)
    mix = session(host_share=0.3)
    b = mix.run().reshape(mix.nsnap, 6, orbits)
    ms, n_host = mix.host_stats()
    assert n_host == 60 and ms > 0
    n_gpu = orbits - n_host
    np.testing.assert_array_equal(b[:, :, :n_gpu], a[:, :, :n_gpu])
    ref = Oracle("oculomotor").run({0: "em", 2: "rk4", 3: "ie"}[solver], cfg["y0"], params, dt, tspan, ks,
                                   counters=counters, nnoi=nnoi)
    assert_traj_close(b, ref["out"], what="host share")
    if solver == 0:
        np.testing.assert_array_equal(mix.counters(), ref["counters"])
    if solver == 3:
        np.testing.assert_array_equal(mix.newton_updates(), ref["newton_updates"])
    # without trajectory output the host share integrates the same orbits
    if solver in (0, 3):
        none = si.Session(cfg["system"], solver, orbits, 6, npar, nnoi, dt, tspan, ks, 0,
                          dict(schedule="lane", host_share=0.3), output_type=0)
        none.setup(cfg["y0"], params, None, counters)
        assert none.run().size == 0
        assert none.host_stats()[1] == n_host
        if solver == 0:
            np.testing.assert_array_equal(none.counters(), ref["counters"])
        else:
            np.testing.assert_array_equal(none.newton_updates(), ref["newton_updates"])
    # the second run starts from where both shares ended
    b2 = mix.run().reshape(mix.nsnap, 6, orbits)
    np.testing.assert_array_equal(b2[0], b[-1])
    a2 = gpu.run().reshape(gpu.nsnap, 6, orbits)
    np.testing.assert_array_equal(b2[:, :, :n_gpu], a2[:, :, :n_gpu])
    assert_traj_close(b2[:, :, n_gpu:], a2[:, :, n_gpu:], what="continued host share")
