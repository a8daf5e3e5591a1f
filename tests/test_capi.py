"""CPU tests of the native library: exported symbols, host logic, program
assembly and offline gfx950 compilation of every solver and schedule.
No GPU is touched (no kernel is launched)."""
import ctypes
import os
import re

import numpy as np
import pytest

import sodecl_amd
from sodecl_amd import sodecl_interface as si
from oracle.oracle import num_launches as oracle_num_launches

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sodecl.h")
REF = "/root/reference"


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(sodecl_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(sodecl_amd.library_path())
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_num_launches_matches_oracle_rule():
    for tspan, dt, ks in [(0.1, 1e-6, 400), (0.1, 1e-8, 40000), (20.0, 1e-3, 100), (10.0, 0.01, 10),
                          (1.0, 0.3, 1), (0.0, 0.1, 1), (50, 0.05, 40), (0.3, 0.1, 1)]:
        assert si.num_launches(tspan, dt, ks) == oracle_num_launches(tspan, dt, ks)


def test_program_carries_reference_macro_contract():
    s = si.Session(sodecl_amd.system_path("oculomotor"), 3, 17, 6, 6, 0, 1.23456789e-4, 1.0, 7, 0, {})
    s.compile_only("gfx950")
    src = s.program_source()
    for line in ["#define _numnoi_ 0", "#define _listsize_ 17", "#define _numsteps_ 7", "#define _m_dt_ 0.000123457",
                 "#define _numeq_ 6", "#define _numpar_ 6", "#define _epsilon1_ 1e-07"]:
        assert line in src, line
    assert src.index("#define _numnoi_") < src.index("void sode_system(") < src.index("\nsodecl_lane(")


def test_feature_detection_ignores_comments():
    src = "/* sode_system_i( */\n// calc_jacobian_band(\nvoid sode_system(double t, double y[], double yo[], double p[]) {}\n"
    s = si.Session(src, 1, 4, 1, 1, 0, 0.1, 1.0, 1, 0, {"schedule": "group"}, True)
    s.compile_only("gfx950")
    assert "#define SODECL_HAS_COMPONENT 0" in s.program_source()
    assert "#define SODECL_HAS_JACOBIAN_BAND 0" in s.program_source()


CASES = [
    ("oculomotor", 1, 6, 6, 0, "lane"), ("oculomotor", 2, 6, 6, 0, "lane"), ("oculomotor", 3, 6, 6, 0, "lane"),
    ("oculomotor", 4, 6, 6, 0, "lane"), ("oculomotor", 0, 6, 8, 2, "lane"),
    ("kuramoto", 1, 10, 11, 0, "group"), ("kuramoto", 2, 64, 65, 0, "group"), ("kuramoto", 0, 32, 65, 32, "group"),
    ("kuramoto", 3, 8, 9, 0, "lane"), ("kuramoto", 4, 8, 9, 0, "lane"), ("kuramoto", 2, 10, 11, 0, "lane"),
    ("brusselator", 3, 32, 3, 0, "lane"), ("brusselator", 4, 16, 3, 0, "lane"),
]


@pytest.mark.parametrize("system,solver,neq,npar,nnoi,schedule", CASES)
def test_offline_compile_gfx950(system, solver, neq, npar, nnoi, schedule):
    s = si.Session(sodecl_amd.system_path(system), solver, 1000, neq, npar, nnoi, 1e-3, 1.0, 10, 0,
                   {"schedule": schedule})
    assert s.compile_only("gfx950") > 1000


@pytest.mark.parametrize("solver", [3, 4])
def test_offline_compile_batch(solver):
    b = si.Batch(sodecl_amd.system_path("brusselator"), solver, 0.01, 50, 3, [32, 128, 512], 0,
                 {"newton_max_iter": 5})
    assert b.compile_only("gfx950") > 1000
    src = b.program_source()
    assert len(re.findall(r"namespace sodecl_cls\d+ \{", src)) == 3


def test_build_error_is_reported():
    bad = "void sode_system(double t, double y[_numeq_], double yout[_numeq_], double p[_numpar_]) { yout[0] = ; }"
    s = si.Session(bad, 1, 4, 1, 1, 0, 0.1, 1.0, 1, 0, {}, True)
    with pytest.raises(RuntimeError, match="Failed to build"):
        s.compile_only("gfx950")


def test_invalid_configuration_rejected():
    with pytest.raises(RuntimeError, match="solver"):
        si.Session(sodecl_amd.system_path("oculomotor"), 9, 4, 6, 6, 0, 0.1, 1.0, 1, 0, {})
    with pytest.raises(RuntimeError):
        si.Session(sodecl_amd.system_path("oculomotor"), 1, 0, 6, 6, 0, 0.1, 1.0, 1, 0, {})
    with pytest.raises(RuntimeError, match="opening"):
        si.Session("/nonexistent/system.cl", 1, 4, 6, 6, 0, 0.1, 1.0, 1, 0, {})
    with pytest.raises(ValueError, match="unknown option"):
        si.Session(sodecl_amd.system_path("oculomotor"), 1, 4, 6, 6, 0, 0.1, 1.0, 1, 0, {"bogus": 1})


def test_invalid_solver_in_sodeclcall_returns_empty():
    out = si.sodeclcall(np.zeros(2), np.zeros(12), np.ones(12), 0, 0, sodecl_amd.system_path("oculomotor"), 7, 2, 6,
                        6, 0, 1e-3, 0.01, 10, 0)
    assert out.size == 0


def test_data_file_readers(tmp_path):
    """read_binary_data_from_file / read_data_from_file (sodeclmgr.cpp:839-914)."""
    lib = ctypes.CDLL(sodecl_amd.library_path())
    for name in ("sodecl_read_binary_data_from_file", "sodecl_read_data_from_file"):
        getattr(lib, name).argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]
    vals = np.array([1.5, -2.25, 3e-7, 4e10])
    (tmp_path / "d.bin").write_bytes(vals.tobytes())
    (tmp_path / "d.txt").write_text("1.5 -2.25\n\n3e-7   4e10\n")
    for fname, fn in (("d.bin", lib.sodecl_read_binary_data_from_file), ("d.txt", lib.sodecl_read_data_from_file)):
        out = np.zeros(8)
        n = ctypes.c_size_t()
        assert fn(str(tmp_path / fname).encode(), out.ctypes.data, 8, ctypes.byref(n)) == 1
        assert n.value == 4
        np.testing.assert_array_equal(out[:4], vals)
    out = np.zeros(2)
    n = ctypes.c_size_t()
    assert lib.sodecl_read_data_from_file(str(tmp_path / "missing.txt").encode(), out.ctypes.data, 2,
                                          ctypes.byref(n)) == 0


def test_adjoint_workspace_and_checkpoints():
    s = si.Session(sodecl_amd.system_path("kuramoto"), 2, 16384, 64, 65, 0, 0.01, 20.0, 10, 0,
                   {"schedule": "group"})
    assert s.default_checkpoint_interval() == 45          # ceil(sqrt(2000))
    assert s.checkpoint_count(45) == 45                   # ceil(2000 / 45)
    assert s.adjoint_workspace_size(45) == 16384 * 45 * (4 * 64 + 1)
    lane = si.Session(sodecl_amd.system_path("oculomotor"), 2, 16, 6, 6, 0, 0.01, 1.0, 10, 0, {})
    assert lane.adjoint_workspace_size(10) == 0


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="reference tree not present")
@pytest.mark.parametrize("name,solver,neq,npar,nnoi", [("oculomotor.cl", 2, 6, 6, 0), ("oculomotor.cl", 3, 6, 6, 0),
                                                       ("oculomotor.cl", 0, 6, 8, 2), ("kuramoto.cl", 1, 5, 6, 0),
                                                       ("kuramoto.cl", 0, 5, 11, 5)])
def test_reference_example_plugins_compile_unchanged(name, solver, neq, npar, nnoi):
    """The reference's own system sources go through the JIT as they are."""
    s = si.Session(os.path.join(REF, "examples", name), solver, 64, neq, npar, nnoi, 1e-3, 1.0, 10, 0, {})
    assert s.compile_only("gfx950") > 1000


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "interfaces", "python", "sodecl.cpp")),
                    reason="reference tree not present")
def test_reference_pybind_binding_builds_against_sodeclmgr(tmp_path):
    """The reference's own binding (interfaces/python/sodecl.cpp), with only
    its include switched to sodeclmgr.hpp, compiles and links against this
    library: the sodeclmgr surface it uses is all there."""
    import subprocess
    import sysconfig

    import pybind11
    src = open(os.path.join(REF, "interfaces", "python", "sodecl.cpp")).read()
    src = src.replace('#include "sodecl.hpp"', '#include "sodeclmgr.hpp"')
    cpp = tmp_path / "sodecl.cpp"
    cpp.write_text(src)
    out = tmp_path / ("sodecl_interface" + sysconfig.get_config_var("EXT_SUFFIX"))
    pkg = os.path.join(ROOT, "sodecl_amd")
    cmd = ["g++", "-O0", "-shared", "-fPIC", "-std=c++17", f"-I{pkg}/csrc", f"-I{ROOT}/include",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", str(cpp),
           os.path.join(pkg, "csrc", "sodeclmgr.cpp"), f"-L{pkg}", "-lsodecl_amd", f"-Wl,-rpath,{pkg}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert out.exists()


def test_pairwise_form_selected_and_lds_bounded():
    """Kuramoto defines sode_pair/sode_system_pairsum: the group schedule uses
    the pairwise form automatically: the full N x (N+1) LDS matrix when the
    group has at least two lanes per component and the matrix is at most
    16 KiB per orbit, by lane exchange (layout 3, no LDS) when it has one,
    else the full matrix while it fits in 16 KiB; half storage, the full
    matrix and the lane exchange when asked for, never when disabled;
    blocks stay small enough that their LDS storage fits in 48 KiB."""
    path = sodecl_amd.system_path("kuramoto")

    def layout(n, g, mode):
        s = si.Session(path, 2, 1000, n, n + 1, 0, 1e-3, 1.0, 10, 0,
                       {"schedule": "group", "group_size": g, "pairwise": mode})
        assert s.compile_only("gfx950") > 1000
        src = s.program_source()
        value = int(src.split("#define SODECL_USE_PAIRS ")[1].split()[0])
        sched, group, block = s.launch_config()
        assert group == g
        per_orbit = {0: 0, 1: n * (n + 1) * 8, 2: n * ((n - 1) // 2 + (n % 2 == 0)) * 8, 3: 0}[value]
        assert (block // group) * per_orbit <= 48 * 1024 or block == group
        return value

    # (N, G, automatic layout, layout for "always", layout for "exchange")
    for n, g, auto, always, exch in ((64, 64, 3, 3, 3), (32, 32, 3, 3, 3), (10, 16, 3, 3, 3), (45, 64, 3, 3, 3),
                                     (64, 16, 0, 2, 0), (32, 16, 1, 1, 0), (44, 32, 1, 1, 0), (45, 32, 0, 2, 0),
                                     (10, 32, 1, 1, 3), (10, 64, 1, 1, 3), (32, 64, 1, 1, 3)):
        assert layout(n, g, 0) == auto, (n, g)
        assert layout(n, g, 1) == always, (n, g)
        assert layout(n, g, 4) == exch, (n, g)
        assert layout(n, g, -1) == 0
        assert layout(n, g, 2) == 2
        assert layout(n, g, 3) == 1
    lane = si.Session(path, 2, 1000, 10, 11, 0, 1e-3, 1.0, 10, 0, {"schedule": "lane"})
    lane.compile_only("gfx950")
    assert "#define SODECL_USE_PAIRS 0" in lane.program_source()


def test_group_blocks_fill_the_chip():
    """Group-schedule blocks shrink (down to one wavefront) while a launch
    would give fewer blocks than the 256 compute units; an explicit block
    size is kept.  With the pairwise drift a launch of less than a wavefront
    per SIMD first widens its groups."""
    path = sodecl_amd.system_path("kuramoto")

    def cfg(orbits, n, **opts):
        s = si.Session(path, 1, orbits, n, n + 1, 0, 1e-2, 1.0, 10, 0, dict(schedule="group", **opts))
        s.compile_only("gfx950")
        return s.launch_config()

    assert cfg(1024, 10, group_size=16) == (2, 16, 64)
    assert cfg(2048, 10, group_size=16) == (2, 16, 128)
    assert cfg(4096, 10) == (2, 16, 256)
    assert cfg(16384, 64) == (2, 64, 256)
    assert cfg(1024, 10, group_size=16, block_size=256) == (2, 16, 256)
    assert cfg(16, 64) == (2, 64, 64)
    # with the pairwise drift, small launches widen the group (up to a
    # wavefront per orbit) so more lanes share the pair evaluations
    assert cfg(1024, 10) == (2, 64, 256)
    assert cfg(2048, 10) == (2, 32, 256)
    assert cfg(1024, 10, pairwise=-1) == (2, 16, 64)


def test_team_schedule_selection():
    """The team schedule needs the plugin's sode_system_team (oculomotor has
    it, Kuramoto does not) and an explicit non-stochastic solver.  It is
    never chosen automatically (slower than the lane schedule for the
    oculomotor model: the burst curves are only part of the step)."""
    ocu = sodecl_amd.system_path("oculomotor")
    for solver in (1, 2):
        s = si.Session(ocu, solver, 65536, 6, 6, 0, 1e-3, 1.0, 10, 0, {"schedule": "team"})
        assert s.compile_only("gfx950") > 1000
        assert s.launch_config() == (4, 2, 128)
        assert "\nsodecl_team(" in s.program_source()
    auto = si.Session(ocu, 2, 65536, 6, 6, 0, 1e-3, 1.0, 10, 0, {})
    auto.compile_only("gfx950")
    assert auto.launch_config()[0] == 1
    for solver, nnoi, npar in ((3, 0, 6), (4, 0, 6), (0, 2, 8)):
        s = si.Session(ocu, solver, 300, 6, npar, nnoi, 1e-3, 1.0, 10, 0, {"schedule": "team"})
        s.compile_only("gfx950")
        assert s.launch_config()[0] == 1
    k = si.Session(sodecl_amd.system_path("kuramoto"), 2, 300, 6, 7, 0, 1e-3, 1.0, 10, 0, {"schedule": "team"})
    k.compile_only("gfx950")
    assert k.launch_config()[0] == 1
    t4 = si.Session(ocu, 2, 300, 6, 6, 0, 1e-3, 1.0, 10, 0, {"schedule": "team", "group_size": 4,
                                                            "block_size": 256})
    t4.compile_only("gfx950")
    assert t4.launch_config() == (4, 4, 256)


def test_occupancy_hint_and_lane_block():
    """At most one wavefront per SIMD (1024 SIMDs assumed without a GPU):
    the kernels carry the one-wave occupancy hint and the lane schedule keeps
    128-lane blocks; above that, no hint and 256-lane blocks."""
    ocu = sodecl_amd.system_path("oculomotor")
    small = si.Session(ocu, 2, 65536, 6, 6, 0, 1e-3, 1.0, 10, 0, {})
    small.compile_only("gfx950")
    assert "amdgpu_waves_per_eu(1, 1)" in small.program_source()
    assert small.launch_config() == (1, 1, 128)
    big = si.Session(ocu, 2, 65537, 6, 6, 0, 1e-3, 1.0, 10, 0, {})
    big.compile_only("gfx950")
    assert "#define SODECL_KERNEL_ATTR __attribute__" not in big.program_source()
    assert big.launch_config() == (1, 1, 256)
    # group schedule: lanes = orbits x group size
    kur = sodecl_amd.system_path("kuramoto")
    g = si.Session(kur, 2, 16384, 64, 65, 0, 1e-3, 1.0, 10, 0, {"schedule": "group"})
    g.compile_only("gfx950")
    assert "#define SODECL_KERNEL_ATTR __attribute__" not in g.program_source()
    g = si.Session(kur, 1, 1024, 10, 11, 0, 1e-3, 1.0, 10, 0, {"schedule": "group"})
    g.compile_only("gfx950")
    assert "#define SODECL_KERNEL_ATTR __attribute__" in g.program_source()


def test_sampler_rejects_bad_tables_before_launch():
    """Validation happens on the host before any launch (no GPU needed)."""
    with pytest.raises(RuntimeError):
        si.sample_device([(2, 0.0, 1.0)], 10, 1, 1 << 20, 0)    # log-uniform needs a, b > 0
    with pytest.raises(RuntimeError):
        si.sample_device([(7, 0.0, 1.0)], 10, 1, 1 << 20, 0)    # unknown kind
    with pytest.raises(RuntimeError):
        si.sample_device([], 10, 1, 1 << 20, 0)                 # no columns


def _host_share_library(system, solver, dt, tspan, ksteps, neq, npar, nnoi, orbits):
    lib = ctypes.CDLL(sodecl_amd.library_path())
    lib.sodecl_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6
    lib.sodecl_host_library.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.sodecl_host_library.restype = ctypes.c_size_t
    lib.sodecl_last_error.argtypes = [ctypes.c_void_p]
    lib.sodecl_last_error.restype = ctypes.c_char_p
    lib.sodecl_destroy.argtypes = [ctypes.c_void_p]
    ctx = ctypes.c_void_p()
    assert lib.sodecl_create(ctypes.byref(ctx), b"kernels", sodecl_amd.system_path(system).encode(), solver, dt, tspan,
                             ksteps, neq, npar, nnoi, orbits, 1) == 1
    buf = ctypes.create_string_buffer(4096)
    n = lib.sodecl_host_library(ctx, buf, len(buf))
    err = lib.sodecl_last_error(ctx)
    lib.sodecl_destroy(ctx)
    assert n > 0, err
    host = ctypes.CDLL(buf.value.decode())
    f = host.sodecl_host_lane
    f.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int] * 6 + [ctypes.c_longlong, ctypes.c_int]
    return f


@pytest.mark.parametrize("solver,name", [(2, "rk4"), (1, "euler"), (3, "ie"), (4, "im"), (0, "em")])
def test_host_share_library_matches_oracle(solver, name):
    """The host share's library (the product's lane integrator and the
    system source compiled for the host, built here through the C ABI)
    integrates an orbit range of host arrays: within 1e-10 of the oracle,
    Threefry counters and Newton counts exact, orbits outside the range
    untouched."""
    from oracle.oracle import Oracle
    from sodecl_amd import configs
    from tests.parity import assert_traj_close
    orbits, first = 40, 13
    explicit = solver in (0, 1, 2)
    dt, tspan, ks = (2e-6, 0.01, 100) if explicit else (1e-5, 0.02, 50)
    cfg = configs.oculomotor_rk4(orbits=orbits, seed=5, dt=dt, tspan=tspan, ksteps=ks)
    cfg["y0"][:, 5] += 1.0
    npar, nnoi = 6, 0
    params = np.ascontiguousarray(cfg["params"])
    counters = None
    if solver == 0:
        nnoi, npar = 2, 8
        rng = np.random.default_rng(5)
        params = np.ascontiguousarray(np.concatenate([params, rng.uniform(0, 0.2, (orbits, 2))], 1))
        counters = rng.integers(0, 10 ** 11, orbits).astype(np.float64)
    f = _host_share_library("oculomotor", solver, dt, tspan, ks, 6, npar, nnoi, orbits)
    launches = oracle_num_launches(tspan, dt, ks)
    t = np.zeros(orbits)
    y = np.ascontiguousarray(cfg["y0"]).copy()
    cnt = None if counters is None else counters.copy()
    out = np.full((launches + 1, 6, orbits), np.nan)
    nu = np.zeros(orbits, dtype=np.int64)
    f(t.ctypes.data, y.ctypes.data, params.ctypes.data, None if cnt is None else cnt.ctypes.data, out.ctypes.data,
      nu.ctypes.data, orbits, first, orbits, 0, launches, 1, 0, 4)
    ref = Oracle("oculomotor").run(name, cfg["y0"], params, dt, tspan, ks, counters=counters, nnoi=nnoi)
    assert_traj_close(out[:, :, first:], ref["out"][:, :, first:], what=f"host share {name}")
    assert np.isnan(out[:, :, :first]).all()
    np.testing.assert_array_equal(y[:first], cfg["y0"][:first])
    assert_traj_close(y[first:].T[None], ref["y"][first:].T[None], what="final state")
    np.testing.assert_array_equal(t[first:], ref["t"][first:])
    if solver == 0:
        np.testing.assert_array_equal(cnt[first:], ref["counters"][first:])
        np.testing.assert_array_equal(cnt[:first], counters[:first])
    if not explicit:
        np.testing.assert_array_equal(nu[first:], ref["newton_updates"][first:])


def test_jit_options_reach_the_compiler_and_the_cache_key(tmp_path, monkeypatch):
    """SODECL_AMD_JIT_OPTIONS is handed to hipRTC (an unknown option fails
    the build) and is part of the code-object cache key (another option set
    is compiled and stored beside the first); a scheduling strategy among
    the options replaces the ILP strategy of one-wavefront launches instead
    of being given a second time."""
    monkeypatch.setenv("SODECL_AMD_CACHE_DIR", str(tmp_path))

    def compile_with(opts):
        if opts is None:
            monkeypatch.delenv("SODECL_AMD_JIT_OPTIONS", raising=False)
        else:
            monkeypatch.setenv("SODECL_AMD_JIT_OPTIONS", opts)
        s = si.Session(sodecl_amd.system_path("oculomotor"), 2, 65536, 6, 6, 0, 1e-3, 1.0, 100, 0,
                       {"schedule": "lane"})
        return s.compile_only("gfx950")

    def stored():
        return sorted(p.name for p in tmp_path.iterdir() if p.suffix == ".hsaco")

    assert compile_with(None) > 1000
    before = stored()
    assert compile_with("-mllvm -amdgpu-sched-strategy=iterative-ilp") > 1000
    after = stored()
    assert len(after) == len(before) + 1
    with pytest.raises(RuntimeError, match="unknown argument: '--sodecl-no-such-option'"):
        compile_with("--sodecl-no-such-option")
