"""Pins the CPU oracle (oracle/sodecl_oracle.c) before anything is checked
against it.

Checks, each chosen so that a plausible mistake fails one of them:
* the reference's own golden results (tests/data/oculomotor_*.pkl, decoded to
  tests/golden/*.npy without unpickling) for Euler, RK4, implicit Euler and
  implicit midpoint -- these pin the step formulas, the quirky tridiagonal
  solve, the Newton stopping rules, the launch count, the snapshot rule and
  the [snapshot][equation][orbit] output order;
* Random123's published Threefry-2x64-20 known-answer vectors;
* closed forms: u01 end points, Box-Muller radius, Kuramoto N=2, the
  6-significant-digit kernel dt;
* the reference kernels compiled as C (oracle/_ref) where they were built.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, RefKernels, have_ref, num_launches

GOLD = {
    # name: (solver, dt, tspan, ksteps)  -- tests/sodecl_python_tests.py:9-179
    "euler": ("euler", 1e-8, 0.1, 40000),
    "rk": ("rk4", 1e-6, 0.1, 400),
    "ie": ("ie", 1e-6, 0.1, 400),
    "im": ("im", 1e-6, 0.1, 400),
}
OCU_Y0 = [0, 0, 0, 0, 0, 2.0]
OCU_P = [120, 1.5, 0.0045, 0.05, 600, 9]


def fixture(name):
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return np.load(os.path.join(here, "golden", f"oculomotor_{name}.npy"))


@pytest.fixture(scope="module")
def ocu():
    return Oracle("oculomotor")


@pytest.fixture(scope="module")
def kura():
    return Oracle("kuramoto")


@pytest.mark.parametrize("name", list(GOLD))
def test_oculomotor_fixture(ocu, name):
    solver, dt, tspan, ksteps = GOLD[name]
    y0 = np.tile(OCU_Y0, (2, 1))
    p = np.tile(OCU_P, (2, 1))
    r = ocu.run(solver, y0, p, dt, tspan, ksteps)
    out = r["out"].reshape(r["out"].shape[0], -1).T      # sodecl.py:79-80 view
    gold = fixture(name)
    assert out.shape == gold.shape
    # the reference's own test: numpy.isclose defaults (sodecl_python_tests.py:50)
    assert np.isclose(out, gold).all()
    # and the much tighter agreement actually achieved
    rel = np.abs(out - gold) / np.maximum(np.abs(gold), 1.0)
    assert rel.max() < 1e-11


@pytest.mark.parametrize("name", ["ie", "im"])
def test_planted_textbook_thomas_fails_fixture(name):
    """A textbook tridiagonal solve planted in the oracle (built as
    liboracle_oculomotor_planted.so) moves the implicit results outside the
    reference's own tolerance, so the fixtures pin the reference's solve."""
    solver, dt, tspan, ksteps = GOLD[name]
    planted = Oracle("oculomotor_planted")
    r = planted.run(solver, np.tile(OCU_Y0, (2, 1)), np.tile(OCU_P, (2, 1)), dt, tspan, ksteps)
    out = r["out"].reshape(r["out"].shape[0], -1).T
    assert not np.isclose(out, fixture(name)).all()


def test_launch_count_rule():
    # 0.1/(1e-6*400) = 250.00000000000003 -> 251 launches (252 snapshots)
    assert num_launches(0.1, 1e-6, 400) == 251
    # 0.1/(1e-8*40000) = 250.0 -> 250 launches
    assert num_launches(0.1, 1e-8, 40000) == 250
    assert num_launches(10.0, 0.01, 10) == 100
    assert num_launches(0.0, 0.01, 10) == 0


def test_kernel_dt_six_digits(ocu):
    assert ocu.kernel_dt(1e-6) == 1e-6
    assert ocu.kernel_dt(0.1234567891) == 0.123457
    assert ocu.kernel_dt(1.0 / 3.0) == 0.333333


def test_threefry_published_kats(kura):
    # Random123 kat_vectors, threefry2x64 with 20 rounds
    assert kura.threefry((0, 0), (0, 0)) == (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)
    ones = 0xFFFFFFFFFFFFFFFF
    assert kura.threefry((ones, ones), (ones, ones)) == (0xE02CB7C4D95D277A, 0xD06633D0893B8B68)
    assert kura.threefry((0x243F6A8885A308D3, 0x13198A2E03707344),
                         (0xA4093822299F31D0, 0x082EFA98EC4E6C89)) == (0x263C7D30BB0F0AF1, 0x56BE8361D3311526)


def test_u01_closed_form(kura):
    # (1 + (x >> 11)) * (2^53 - 1) * 2^-106: strictly inside (0, 1)
    assert kura.u01(0) == (2.0 ** 53 - 1) * 2.0 ** -106
    assert kura.u01(2 ** 64 - 1) == (2.0 ** 53) * (2.0 ** 53 - 1) * 2.0 ** -106
    assert 0.0 < kura.u01(2 ** 64 - 1) < 1.0
    x = 0x03ED3FEB0846F822
    assert kura.u01(x) == float((1 + (x >> 11))) * ((2.0 ** 53 - 1) * 2.0 ** -106)


def test_box_muller_pair(kura):
    """Noise pair = sqrt(dt) * r * (cos, sin) with r^2 = -2 log u0; checks the
    radius (independent of the angle) from the Threefry output."""
    dt = 1e-3
    for orbit, cw in [(0, 0), (1, 7), (123456, 99999)]:
        w0, w1 = kura.threefry((orbit, cw), (0, 0xF00DCAFE))
        u0 = kura.u01(w0)
        u1 = kura.u01(w1)
        pair = kura.noise_pair(orbit, cw, dt)
        r2 = (pair[0] ** 2 + pair[1] ** 2) / dt
        assert r2 == pytest.approx(-2 * math.log(u0), rel=1e-13)
        ang = math.atan2(pair[1], pair[0]) % (2 * math.pi)
        assert ang == pytest.approx((2 * math.pi * u1) % (2 * math.pi), abs=1e-12)


def test_kuramoto_rhs_closed_form(kura):
    y = np.array([0.3, 1.9])
    K, w1, w2 = 2.5, 0.7, 1.3
    f = kura.rhs(y, [K, w1, w2])
    assert f[0] == pytest.approx(w1 + (K / 2) * math.sin(y[1] - y[0]), rel=1e-15)
    assert f[1] == pytest.approx(w2 + (K / 2) * math.sin(y[0] - y[1]), rel=1e-15)
    # brute force N=5 against the definition
    rng = np.random.default_rng(1)
    y = rng.uniform(0, 6, 5)
    p = rng.uniform(0, 2, 6)
    f = kura.rhs(y, p)
    want = [p[i + 1] + p[0] / 5 * sum(math.sin(y[j] - y[i]) for j in range(5)) for i in range(5)]
    np.testing.assert_allclose(f, want, rtol=1e-14)


def test_em_counter_advance(kura):
    rng = np.random.default_rng(2)
    orbits, n = 8, 3
    y0 = rng.uniform(0, 6, (orbits, n))
    p = np.concatenate([rng.uniform(0, 5, (orbits, 1)), rng.normal(1, 0.1, (orbits, n)),
                        rng.uniform(0, 0.5, (orbits, n))], 1)
    c0 = rng.integers(0, 10 ** 11, orbits).astype(np.float64)
    r = kura.run("em", y0, p, 1e-2, 1.0, 10, counters=c0, nnoi=n)
    # ceil(3/2) = 2 Threefry calls per step, 100 steps
    np.testing.assert_array_equal(r["counters"] - c0, 200.0)
    assert np.all(np.isfinite(r["out"]))


def test_em_first_step_by_hand(kura):
    """One EM step of a 2-oscillator Kuramoto built by hand from the
    KAT-verified Threefry: y1 = y0 + f(y0) dt + sigma * noise."""
    dt = 0.01
    y0 = np.array([[0.4, 2.2]])
    p = np.array([[1.5, 0.9, 1.1, 0.2, 0.3]])
    c0 = np.array([5.0])
    r = kura.run("em", y0, p, dt, dt, 1, counters=c0, nnoi=2)
    noise = kura.noise_pair(0, 5, dt)
    f = kura.rhs(y0[0], p[0, :3])
    want = y0[0] + f * dt + p[0, 3:] * noise
    np.testing.assert_allclose(r["out"][1, :, 0], want, rtol=1e-15, atol=0)


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("system,solver", [
    ("oculomotor", "euler"), ("oculomotor", "rk4"), ("oculomotor", "ie"), ("oculomotor", "im"),
    ("kuramoto", "euler"), ("kuramoto", "rk4"), ("kuramoto", "ie"), ("kuramoto", "im"),
    ("brusselator", "ie"), ("brusselator", "im"),
])
def test_port_matches_reference_kernels_ode(system, solver):
    rng = np.random.default_rng(3)
    orbits = 16
    if system == "oculomotor":
        y0 = rng.normal(0, 1e-3, (orbits, 6)); y0[:, 5] += 1.0
        p = np.array(OCU_P) * np.exp(rng.uniform(np.log(0.5), np.log(2), (orbits, 6)))
        dt, tspan, ks = 1e-5, 2e-3, 20
    elif system == "kuramoto":
        n = 7
        y0 = rng.uniform(0, 2 * np.pi, (orbits, n))
        p = np.concatenate([rng.uniform(0, 5, (orbits, 1)), rng.uniform(0.5, 1.5, (orbits, n))], 1)
        dt, tspan, ks = 0.01, 1.0, 10
    else:
        n = 16
        p = np.stack([np.ones(orbits), rng.uniform(2, 4, orbits), rng.uniform(0.01, 1, orbits)], 1)
        x = (np.arange(n // 2) + 1) / (n // 2 + 1)
        y0 = np.zeros((orbits, n))
        y0[:, 0::2] = 1 + 0.5 * np.sin(2 * np.pi * x)
        y0[:, 1::2] = p[:, 1:2]
        dt, tspan, ks = 0.01, 0.5, 5
    a = Oracle(system).run(solver, y0, p, dt, tspan, ks)
    b = RefKernels(system, solver).run(y0, p, dt, tspan, ks)
    np.testing.assert_array_equal(a["out"], b["out"])


@pytest.mark.skipif(not have_ref("kuramoto", "em"), reason="oracle/_ref not built")
@pytest.mark.parametrize("system", ["kuramoto", "refkuramoto"])
def test_port_matches_reference_kernels_em(kura, system):
    rng = np.random.default_rng(4)
    orbits, n = 32, 5
    y0 = rng.uniform(0, 2 * np.pi, (orbits, n))
    p = np.concatenate([rng.uniform(0, 5, (orbits, 1)), rng.normal(1, 0.1, (orbits, n)),
                        rng.uniform(0, 0.5, (orbits, n))], 1)
    c0 = rng.integers(0, 10 ** 11, orbits).astype(np.float64)
    a = kura.run("em", y0, p, 1e-3, 0.5, 50, counters=c0, nnoi=n)
    b = RefKernels(system, "em").run(y0, p, 1e-3, 0.5, 50, counters=c0, nnoi=n)
    np.testing.assert_array_equal(a["out"], b["out"])
    np.testing.assert_array_equal(a["counters"], b["counters"])


@pytest.mark.skipif(not have_ref("refoculomotor", "rk4"), reason="oracle/_ref not built")
@pytest.mark.parametrize("solver", ["euler", "rk4", "ie", "im"])
def test_plugin_matches_reference_example(solver):
    """Our oculomotor plugin against the reference's examples/oculomotor.cl,
    both run through the reference kernels: identical bits."""
    y0 = np.tile(OCU_Y0, (2, 1))
    p = np.tile(OCU_P, (2, 1))
    a = RefKernels("oculomotor", solver).run(y0, p, 1e-6, 2e-3, 40)
    b = RefKernels("refoculomotor", solver).run(y0, p, 1e-6, 2e-3, 40)
    np.testing.assert_array_equal(a["out"], b["out"])
    # randomised parameters and states: both signs of the burst input, the
    # benchmark's parameter spread, blow-ups included (NaN compared as equal)
    from sodecl_amd import configs
    cfg = configs.oculomotor_rk4(orbits=96, seed=11, dt=1e-4, tspan=0.05, ksteps=25)
    y0 = cfg["y0"].copy()
    y0[:, 5] = np.linspace(-2.0, 2.0, 96)
    # zero burst inputs of both signs, and zero states whose derivative
    # terms cancel (the sign of each zero is compared too)
    y0[0, 5], y0[1, 5] = 0.0, -0.0
    y0[2:6] = 0.0
    y0[3, 2] = y0[5, 2] = y0[5, 5] = -0.0
    a = RefKernels("oculomotor", solver).run(y0, cfg["params"], 1e-4, 0.05, 25)
    b = RefKernels("refoculomotor", solver).run(y0, cfg["params"], 1e-4, 0.05, 25)
    np.testing.assert_array_equal(a["out"].view(np.int64), b["out"].view(np.int64))


@pytest.mark.parametrize("d", [6.0, 25.0, 0.15, 9.0, 1.5, 0.0045 * 1.7, -0.3, -7.0, 1e-300])
def test_markstein_division_is_correctly_rounded(ocu, d):
    """The division the kernels use for constant and per-orbit divisors
    (q0 = a/d via the reciprocal, one exact-remainder correction) gives the
    bits of IEEE a / d: random dividends over the normal range plus zeros,
    infinities, NaN and exact quotients.  The select-free form (residual
    clamped at -DBL_MAX) gives the bits of the guarded one wherever d >= 1
    or the dividend's sign bit is set, subnormal and overflowing quotients,
    infinities and NaN included."""
    import ctypes
    f = ocu.lib.orc_markstein_mismatches
    f.restype = ctypes.c_long
    f.argtypes = [ctypes.c_double, ctypes.c_long, ctypes.c_uint64]
    assert f(d, 2_000_000, 12345) == 0


@pytest.mark.parametrize("system,n,npar", [("oculomotor", 6, 6), ("kuramoto", 7, 8)])
def test_plugin_jacobians_match_central_differences(system, n, npar):
    """df/dy (calc_jacobian_exact, else calc_jacobian) and df/dp
    (calc_jacobian_params) of the bundled plugins against central
    differences of sode_system, at states on both sides of the oculomotor
    burst curve's switch."""
    o = Oracle(system)
    rng = np.random.default_rng(3)
    for trial in range(6):
        y = rng.normal(0, 1.0, n)
        if system == "oculomotor":
            y[5] = (-1.0) ** trial * rng.uniform(0.2, 2.0)
            p = np.array(OCU_P) * np.exp(rng.uniform(-0.5, 0.5, 6))
        else:
            p = np.concatenate([[rng.uniform(0, 5)], rng.normal(1, 0.1, n)])
        jac, jp = o.jacobians(y, p)
        for j in range(n):
            h = 1e-6 * max(1.0, abs(y[j]))
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            fd = (o.rhs(yp, p) - o.rhs(ym, p)) / (2 * h)
            np.testing.assert_allclose(jac[:, j], fd, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(fd).max()))
        for k in range(npar):
            h = 1e-6 * max(1.0, abs(p[k]))
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (o.rhs(y, pp) - o.rhs(y, pm)) / (2 * h)
            np.testing.assert_allclose(jp[:, k], fd, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(fd).max()))


def test_oculomotor_newton_matrix_keeps_the_reference_sign():
    """calc_jacobian keeps the reference's J[4][5] (examples/oculomotor.cl:75,
    without the chain rule's minus sign) for the Newton matrix; only the
    exact Jacobian differs, in that entry only."""
    o = Oracle("oculomotor")
    y = np.array([0.1, -0.2, 0.3, 0.4, 0.5, -0.7])
    jn, _ = o.jacobians(y, OCU_P, exact=False)
    je, _ = o.jacobians(y, OCU_P, exact=True)
    diff = np.argwhere(jn != je)
    assert diff.tolist() == [[4, 5]]
    assert je[4, 5] == -jn[4, 5]


@pytest.mark.parametrize("system,solver", [("oculomotor", "euler"), ("oculomotor", "rk4"), ("kuramoto", "euler"),
                                           ("kuramoto", "rk4")])
def test_explicit_sensitivity_matches_central_differences(system, solver):
    """The oracle's forward sensitivity of Euler and RK4 (the exact tangent
    of each step) against central differences of the oracle's own
    integration, on every initial-state and parameter component of three
    orbits, against a random cotangent on every snapshot.  A dropped stage
    term or a transposed Jacobian in the tangent fails it, and so did the
    tangent's earlier use of the oculomotor's Newton matrix (which keeps the
    reference's sign in J[4][5]) in place of df/dy: 8e-4 off for RK4."""
    o = Oracle(system)
    rng = np.random.default_rng(11)
    if system == "oculomotor":
        orbits, n = 3, 6
        y0 = np.tile(OCU_Y0, (orbits, 1)) + rng.normal(0, 0.05, (orbits, n))
        p = np.tile(OCU_P, (orbits, 1)) * np.exp(rng.uniform(-0.3, 0.3, (orbits, 6)))
        # Euler needs the smaller step for these orbits to be well
        # conditioned: at 1e-5 one of them amplifies rounding so much that
        # central differences at h = 1e-6 and 1e-7 disagree in sign
        dt, tspan, ks = (2e-6, 0.01, 250) if solver == "euler" else (1e-5, 0.01, 50)
    else:
        orbits, n = 3, 5
        y0 = rng.uniform(0, 2 * np.pi, (orbits, n))
        p = np.concatenate([rng.uniform(0, 3, (orbits, 1)), rng.normal(1, 0.1, (orbits, n))], 1)
        dt, tspan, ks = 0.01, 1.0, 10
    nsnap = o.run(solver, y0, p, dt, tspan, ks)["out"].shape[0]
    cot = rng.normal(size=(nsnap, n, orbits))
    gy, gp = o.sensitivity_gradient(solver, y0, p, cot, dt, tspan, ks)

    def loss(yv, pv):
        out = o.run(solver, yv, pv, dt, tspan, ks)["out"]
        return np.einsum("sno,sno->o", out, cot)

    for j in range(n):
        h = 1e-6
        yp, ym = y0.copy(), y0.copy()
        yp[:, j] += h
        ym[:, j] -= h
        fd = (loss(yp, p) - loss(ym, p)) / (2 * h)
        np.testing.assert_allclose(gy[:, j], fd, rtol=1e-6, atol=1e-6 * np.abs(gy).max())
    for k in range(p.shape[1]):
        h = 1e-6 * np.abs(p[:, k]).max()
        pp, pm = p.copy(), p.copy()
        pp[:, k] += h
        pm[:, k] -= h
        fd = (loss(y0, pp) - loss(y0, pm)) / (2 * h)
        np.testing.assert_allclose(gp[:, k], fd, rtol=1e-6, atol=1e-6 * np.abs(gp).max())


@pytest.mark.parametrize("system,solver", [("oculomotor", "ie"), ("oculomotor", "im"), ("kuramoto", "ie"),
                                           ("kuramoto", "im")])
def test_implicit_sensitivity_matches_central_differences(system, solver):
    """The oracle's forward sensitivity of the implicit solvers (tangent of
    each step's implicit equation at the converged Newton solution) against
    central differences of the oracle's own integration, run with a Newton
    threshold of 1e-14 so that the truncated iteration is smooth to far
    below the difference step."""
    o = Oracle(system)
    rng = np.random.default_rng(5)
    if system == "oculomotor":
        orbits, n = 3, 6
        y0 = np.tile(OCU_Y0, (orbits, 1)) + rng.normal(0, 0.05, (orbits, n))
        p = np.tile(OCU_P, (orbits, 1)) * np.exp(rng.uniform(-0.3, 0.3, (orbits, 6)))
        dt, tspan, ks = 1e-4, 0.02, 20
    else:
        orbits, n = 3, 5
        y0 = rng.uniform(0, 2 * np.pi, (orbits, n))
        p = np.concatenate([rng.uniform(0, 3, (orbits, 1)), rng.normal(1, 0.1, (orbits, n))], 1)
        dt, tspan, ks = 0.01, 1.0, 10
    nsnap = o.run(solver, y0, p, dt, tspan, ks)["out"].shape[0]
    cot = rng.normal(size=(nsnap, n, orbits))
    gy, gp = o.sensitivity_gradient(solver, y0, p, cot, dt, tspan, ks, newton_tol=1e-14)

    def loss(yv, pv):
        out = o.run(solver, yv, pv, dt, tspan, ks, newton_tol=1e-14)["out"]
        return np.einsum("sno,sno->o", out, cot)

    for j in range(n):
        h = 1e-6
        yp, ym = y0.copy(), y0.copy()
        yp[:, j] += h
        ym[:, j] -= h
        fd = (loss(yp, p) - loss(ym, p)) / (2 * h)
        np.testing.assert_allclose(gy[:, j], fd, rtol=1e-6, atol=1e-6 * np.abs(gy).max())
    for k in range(p.shape[1]):
        h = 1e-6 * np.abs(p[:, k]).max()
        pp, pm = p.copy(), p.copy()
        pp[:, k] += h
        pm[:, k] -= h
        fd = (loss(y0, pp) - loss(y0, pm)) / (2 * h)
        np.testing.assert_allclose(gp[:, k], fd, rtol=1e-6, atol=1e-6 * np.abs(gp).max())


SAMPLE_KEY1 = 0x5EEDD15C0FFEE5ED


def test_sampler_counter_and_key_mapping():
    """Entry (r, c) of the sampler is Threefry-2x64-20 at counter {r, c},
    key {seed, 0x5EEDD15C0FFEE5ED}, converted by u01 (both pinned by their
    known-answer vectors above): uniform(0, 1) returns u0 itself, a constant
    column returns a, and the normal column follows Box-Muller."""
    import math
    o = Oracle("oculomotor")
    seed = 123456789
    table = [(1, 0.0, 1.0), (0, 2.5, 0.0), (3, 1.0, 0.1)]
    out = o.sample(table, 40, seed)
    for r in (0, 1, 17, 39):
        w = o.threefry((r, 0), (seed, SAMPLE_KEY1))
        assert out[r, 0] == o.u01(w[0])
        assert out[r, 1] == 2.5
        w = o.threefry((r, 2), (seed, SAMPLE_KEY1))
        z = math.sqrt(-2.0 * math.log(o.u01(w[0]))) * math.cos(2.0 * math.pi * o.u01(w[1]))
        assert abs(out[r, 2] - (1.0 + 0.1 * z)) <= 4e-16 * 2


def test_sampler_distributions():
    """Moments and ranges of 200000 draws per column (bounds of 5 standard
    errors), and independence of columns with equal distributions."""
    o = Oracle("oculomotor")
    n = 200000
    out = o.sample([(1, 2.0, 5.0), (2, 0.5, 2.0), (3, 1.0, 0.1), (3, 1.0, 0.1)], n, 99)
    u, lu, z1, z2 = out.T
    assert u.min() > 2.0 and u.max() < 5.0
    assert abs(u.mean() - 3.5) < 5 * math.sqrt(0.75 / n)
    assert abs(u.var() - 0.75) < 5 * 0.75 * math.sqrt(0.8 / n)
    assert lu.min() > 0.5 and lu.max() < 2.0
    L = np.log(lu)
    assert abs(L.mean()) < 5 * math.sqrt(math.log(4) ** 2 / 12 / n)
    for z in (z1, z2):
        assert abs(z.mean() - 1.0) < 5 * 0.1 / math.sqrt(n)
        assert abs(z.std() - 0.1) < 5 * 0.1 / math.sqrt(2 * n)
    assert abs(np.corrcoef(z1, z2)[0, 1]) < 5 / math.sqrt(n)


@pytest.mark.skipif(not have_ref("oculomotor", "rk4", simd=True), reason="vectorised reference build not present")
def test_vectorised_reference_build_agrees_with_scalar():
    """The orbit-vectorised build of the reference kernels (the bench's CPU
    baseline, oracle/ref_simd.c) runs the same arithmetic as the scalar
    build up to the vector exponential's rounding: on well-conditioned
    oculomotor orbits (dt = 2e-6) within 1e-12, with a block that is partly
    empty (37 orbits, blocks of 8) and the same snapshot count."""
    rng = np.random.default_rng(21)
    orbits = 37
    y0 = np.tile(OCU_Y0, (orbits, 1)) + rng.normal(0, 1e-3, (orbits, 6))
    p = np.tile(OCU_P, (orbits, 1)) * np.exp(rng.uniform(-0.3, 0.3, (orbits, 6)))
    a = RefKernels("oculomotor", "rk4", fast=True).run(y0, p, 2e-6, 0.01, 100)
    b = RefKernels("oculomotor", "rk4", simd=True).run(y0, p, 2e-6, 0.01, 100)
    assert a["out"].shape == b["out"].shape == (num_launches(0.01, 2e-6, 100) + 1, 6, orbits)
    np.testing.assert_array_equal(a["t"], b["t"])
    scale = np.abs(a["out"]).max(axis=(0, 2), keepdims=True)
    assert np.max(np.abs(a["out"] - b["out"]) / scale) < 1e-12


@pytest.mark.skipif(not have_ref("oculomotorplain", "rk4") or not have_ref("refoculomotor", "rk4"),
                    reason="oracle/_ref not built")
@pytest.mark.parametrize("solver", ["euler", "rk4"])
def test_plain_plugin_matches_reference_example(solver):
    """The plain-style oculomotor plugin (systems/oculomotor_plain.cl, the
    user's way of writing the model) against the reference's
    examples/oculomotor.cl, both through the reference kernels: identical
    bits on the benchmark's parameter spread, burst inputs of both signs and
    blow-ups included."""
    from sodecl_amd import configs
    cfg = configs.oculomotor_rk4(orbits=96, seed=12, dt=1e-4, tspan=0.05, ksteps=25)
    y0 = cfg["y0"].copy()
    y0[:, 5] = np.linspace(-2.0, 2.0, 96)
    y0[0, 5], y0[1, 5] = 0.0, -0.0
    a = RefKernels("oculomotorplain", solver).run(y0, cfg["params"], 1e-4, 0.05, 25)
    b = RefKernels("refoculomotor", solver).run(y0, cfg["params"], 1e-4, 0.05, 25)
    np.testing.assert_array_equal(a["out"].view(np.uint64), b["out"].view(np.uint64))
