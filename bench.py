"""Benchmark: orbit-steps/s of the MI355X integrators on one GPU.

Default workload: the 6-state oculomotor model, classic RK4, dt=1e-3, 20000
steps, a snapshot every 100 steps (201 snapshots written to HBM), 65536
orbits with seeded log-uniform parameters (sodecl_amd/configs.py).  One
"step" of this benchmark is one complete integration of every orbit
(20000 RK4 steps with trajectory writeback), inputs resident in HBM.

Printed (one JSON line on stdout): value = orbits * integrator steps per
run / wall seconds per run; roofline of the integration kernel from HIP
events on its launch stream (FP64 algorithmic flops, 78.6 TFLOP/s peak);
the reference's kernels compiled for the host CPU (oracle/_ref) timed on a
bounded sample on all host cores; and the reference's launch design on the
same GPU (one launch per ksteps segment with a blocking state read-back).
The same line carries the other workloads measured the same way
("workloads": Kuramoto N=10 Euler, 1024 orbits; Euler-Maruyama Kuramoto
N=32, 262144 orbits), the forward+adjoint Kuramoto N=64 RK4 run, the
batched implicit Brusselator runs and the on-device parameter sampling.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
METRIC = "orbit-steps/s (fwd, fwd+adjoint) vs plain kernel, same GPU, at stated tol"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="oculomotor_rk4",
                    choices=["oculomotor_rk4", "kuramoto_euler", "kuramoto_em", "kuramoto_rk4"])
    ap.add_argument("--schedule", default="auto")
    ap.add_argument("--group-size", type=int, default=0)
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-orbits", type=int, default=0, help="orbits in the CPU baseline sample")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched implicit measurement")
    ap.add_argument("--no-sampling", action="store_true", help="skip the on-device parameter sampling measurement")
    ap.add_argument("--no-host-return", action="store_true",
                    help="skip the measurement of sodecl_run returning the trajectory to host memory")
    ap.add_argument("--no-host-share", action="store_true",
                    help="skip the measurement of sodecl_run with and without a host share")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the Kuramoto Euler (1024 orbits) and Euler-Maruyama (262144 orbits) measurements")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run the ahead-of-time streaming copy kernel once over this many MiB "
                         "(reference point for HBM byte counters)")
    return ap.parse_args()


# The committed rocprofv3 summary that states each workload's HBM traffic.
TRAFFIC_PROFILES = {
    "oculomotor_rk4": "oculomotor_rk4_hbm_traffic.md",
    "oculomotor_rk4_plain_plugin": "sodecl_lane_oculomotor_rk4_plain.md",
    "kuramoto_euler": "sodecl_group_kuramoto_euler_n10.md",
    "kuramoto_em": "sodecl_group_kuramoto_em_n32.md",
}


def traffic_from_profiles(name):
    """HBM bytes per launch of the integration kernel as the committed
    rocprofv3 counter summary states them (a line
    'traffic_bytes_per_launch: N', from two --pmc runs, FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size).  This run does not
    measure them: returns (bytes, source) with source naming the file, or
    (None, None)."""
    if name not in TRAFFIC_PROFILES:
        return None, None
    rel = os.path.join("profiles", TRAFFIC_PROFILES[name])
    path = os.path.join(HERE, rel)
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return (float(line.split(":", 1)[1].strip()),
                            f"{rel}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same kernel on this "
                            f"workload (commands in the file), not measured by this run")
    return None, None


def calibration_copy(mib):
    """One launch of the library's streaming copy kernel over mib MiB."""
    import ctypes
    import torch
    import sodecl_amd
    lib = ctypes.CDLL(sodecl_amd.library_path())
    lib.sodecl_copy_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    n = mib * (1 << 20)
    src = torch.ones(n // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    assert lib.sodecl_copy_device(dst.data_ptr(), src.data_ptr(), n, None) == 1
    torch.cuda.synchronize()
    assert torch.equal(src, dst)


def measure(cfg, steps, warmup, opts, label):
    """Times `steps` complete integrations of the workload with inputs
    resident in HBM (state, t and counters reset on the device before each),
    wall clock around all of them and HIP events on the launch stream around
    each integration kernel.  Returns the measurement and the session."""
    import numpy as np
    import torch
    from sodecl_amd import configs, sodecl_interface as si
    orbits, neq = cfg["orbits"], cfg["neq"]
    launches = si.num_launches(cfg["tspan"], cfg["dt"], cfg["ksteps"])
    nsnap = launches + 1
    steps_per_run = launches * cfg["ksteps"]
    dev = torch.device("cuda:0")
    log(f"{label}: compiling and uploading")
    sess = configs.session_for(cfg, **opts)
    sess.setup(cfg["y0"], cfg["params"], None, cfg["counters"])
    y_init = torch.tensor(cfg["y0"], dtype=torch.float64, device=dev).contiguous()
    p = torch.tensor(cfg["params"], dtype=torch.float64, device=dev).contiguous()
    y = torch.empty_like(y_init)
    t = torch.zeros(orbits, dtype=torch.float64, device=dev)
    cnt_init = None if cfg["counters"] is None else torch.tensor(cfg["counters"], dtype=torch.float64, device=dev)
    cnt = None if cnt_init is None else torch.empty_like(cnt_init)
    traj = torch.empty((nsnap, neq, orbits), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    ev = []

    def one_step(record):
        with torch.cuda.stream(stream):
            y.copy_(y_init)
            t.zero_()
            if cnt is not None:
                cnt.copy_(cnt_init)
            if record:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            sess.run_device(t.data_ptr(), y.data_ptr(), p.data_ptr(), 0 if cnt is None else cnt.data_ptr(),
                            traj.data_ptr(), stream.cuda_stream)
            if record:
                e1.record(stream)
                ev.append((e0, e1))

    for i in range(warmup):
        one_step(False)
        torch.cuda.synchronize()
        log(f"{label}: warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        one_step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    log(f"{label}: timed {steps} steps in {wall:.3f} s")
    kms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    flops = cfg["flops_per_orbit_step"] * orbits * steps_per_run
    achieved = flops / (kms * 1e-3) / 1e12
    traffic, traffic_source = traffic_from_profiles(cfg["name"])
    sched = sess.launch_config()
    finite = float(np.isfinite(traj[-1].cpu().numpy()).mean())
    res = dict(
        value=orbits * steps_per_run / (wall / steps), ms_per_step=wall * 1e3 / steps, steps_per_run=steps_per_run,
        nsnap=nsnap, finite=finite,
        config={
            "workload": cfg["name"],
            "system": os.path.basename(cfg["system"]),
            "solver": {0: "euler-maruyama", 1: "euler", 2: "rk4", 3: "implicit-euler", 4: "implicit-midpoint"}[
                cfg["solver"]],
            "orbits": orbits, "neq": neq, "dt": cfg["dt"], "steps": steps_per_run, "ksteps": cfg["ksteps"],
            "snapshots": nsnap,
            "schedule": {1: "lane", 2: "group", 4: "team"}.get(sched[0], str(sched[0])), "group_size": sched[1],
            "block_size": sched[2],
        },
        roofline={
            "bound": "fp64-valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "kernel_ms": kms,
            "flops_per_orbit_step": cfg["flops_per_orbit_step"],
        })
    return res, sess, traj


def cpu_baseline(cfg, orbits_sample):
    """The reference kernels compiled for the host (oracle/_ref), run as the
    reference's host loop does, OpenMP over orbits on every host core: the
    scalar -O3 build, and where one exists the orbit-vectorised build
    (oracle/ref_simd.c: blocks of orbits in SIMD lanes, as an OpenCL CPU
    runtime packs work-items).  The faster of the two is the baseline."""
    from oracle.oracle import RefKernels, have_ref
    system = os.path.splitext(os.path.basename(cfg["system"]))[0]
    solver = {0: "em", 1: "euler", 2: "rk4", 3: "ie", 4: "im"}[cfg["solver"]]
    if not have_ref(system, solver):
        return None
    k = min(orbits_sample, cfg["orbits"])
    nthreads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp:
        nthreads = min(nthreads, int(omp))
    from sodecl_amd import sodecl_interface as si
    steps = si.num_launches(cfg["tspan"], cfg["dt"], cfg["ksteps"]) * cfg["ksteps"]

    def timed(ref):
        t0 = time.perf_counter()
        ref.run(cfg["y0"][:k], cfg["params"][:k], cfg["dt"], cfg["tspan"], cfg["ksteps"],
                counters=None if cfg["counters"] is None else cfg["counters"][:k], nnoi=cfg["nnoi"],
                nthreads=nthreads)
        return time.perf_counter() - t0

    variants = {"scalar": timed(RefKernels(system, solver, fast=True))}
    if cfg["nnoi"] == 0 and have_ref(system, solver, simd=True):
        variants["orbit_vectorised"] = timed(RefKernels(system, solver, simd=True))
    best = min(variants, key=variants.get)
    dt = variants[best]
    return dict(value=k * steps / dt, unit="orbit-steps/s", cores=nthreads, kind="reference",
                variant=best, sample=f"{k} of {cfg['orbits']} orbits, all {steps} steps with trajectory read-back "
                                     f"({dt:.2f} s wall)",
                seconds=dt, host=host_cpu_model(),
                variants={name: dict(value=k * steps / t, seconds=t) for name, t in variants.items()})


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def fwd_adjoint_bench(steps, store_stages=False):
    """Kuramoto N=64, RK4, dt=0.01, 2000 steps, 16384 orbits, a snapshot
    every 10 steps; loss = mean sin/cos misfit against the trajectory of
    seeded ground-truth (K, w); gradients w.r.t. K, w and theta0 through the
    discrete adjoint with ceil(sqrt(2000)) = 45-step checkpoints, or with
    store_stages every step's stage vectors kept by the forward pass (no
    recomputation)."""
    import torch
    from sodecl_amd import configs
    from sodecl_amd.torch_ops import Integrator
    cfg = configs.kuramoto_rk4()
    dev = torch.device("cuda:0")
    integ = Integrator(cfg["system"], cfg["solver"], cfg["orbits"], cfg["neq"], cfg["npar"], cfg["dt"],
                       cfg["tspan"], cfg["ksteps"], store_stages=store_stages)
    y0 = torch.tensor(cfg["y0"], device=dev)
    truth = torch.tensor(cfg["params"], device=dev)
    rng = torch.Generator(device="cpu").manual_seed(1)
    p0 = truth.clone()
    p0[:, 0] *= 1.1
    p0[:, 1:] += (0.02 * torch.randn(p0[:, 1:].shape, generator=rng, dtype=torch.float64)).to(dev)
    with torch.no_grad():
        target = integ(y0, truth)
    st = torch.sin(target)
    ct = torch.cos(target)

    def fwd():
        with torch.no_grad():
            traj = integ(y0, p0)
            return ((torch.sin(traj) - st) ** 2 + (torch.cos(traj) - ct) ** 2).mean()

    def fwd_adj():
        y = y0.clone().requires_grad_(True)
        p = p0.clone().requires_grad_(True)
        traj = integ(y, p)
        loss = ((torch.sin(traj) - st) ** 2 + (torch.cos(traj) - ct) ** 2).mean()
        loss.backward()
        return loss, y.grad, p.grad

    fwd()
    fwd_adj()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fwd()
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, gy, gp = fwd_adj()
    torch.cuda.synchronize()
    t_all = (time.perf_counter() - t0) / steps
    total = cfg["orbits"] * integ.session.launches * cfg["ksteps"]
    if store_stages:
        return dict(workload="kuramoto_rk4_adjoint_stored", orbits=cfg["orbits"], neq=cfg["neq"],
                    steps=total // cfg["orbits"], workspace_GB=integ.session.stored_workspace_size() * 8 / 1e9,
                    fwd_adjoint_value=total / t_all, unit="orbit-steps/s", fwd_adjoint_ms=t_all * 1e3,
                    loss=float(loss.detach()), grad_finite=bool(torch.isfinite(gp).all() and torch.isfinite(gy).all()))
    return dict(workload="kuramoto_rk4_adjoint", orbits=cfg["orbits"], neq=cfg["neq"], steps=total // cfg["orbits"],
                checkpoints=integ.session.checkpoint_count(integ.ckpt_every), ckpt_every=integ.ckpt_every,
                fwd_value=total / t_fwd, fwd_adjoint_value=total / t_all, unit="orbit-steps/s",
                fwd_ms=t_fwd * 1e3, fwd_adjoint_ms=t_all * 1e3, cost_ratio=t_all / t_fwd,
                loss=float(loss.detach()), grad_finite=bool(torch.isfinite(gp).all() and torch.isfinite(gy).all()))


def sampling_bench(reps=5):
    """On-device draw of the Euler-Maruyama Kuramoto parameter table
    (262144 orbits x 65 columns: K ~ U(0,5), w_i ~ N(1,0.1), sigma_i ~
    U(0,0.5)) against numpy's PCG64 draw of the same shape on the host."""
    import numpy as np
    import torch
    from sodecl_amd.sampling import sample
    orbits, n = 262144, 32
    dists = [("uniform", 0, 5)] + [("normal", 1.0, 0.1)] * n + [("uniform", 0, 0.5)] * n
    out = torch.empty((orbits, len(dists)), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        sample(dists, orbits, 1, out=out)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for r in range(reps):
            sample(dists, orbits, r + 2, out=out)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    K = rng.uniform(0, 5, (orbits, 1))
    w = rng.normal(1.0, 0.1, (orbits, n))
    s = rng.uniform(0, 0.5, (orbits, n))
    host = np.concatenate([K, w, s], 1)
    host_ms = (time.perf_counter() - t0) * 1e3
    values = orbits * len(dists)
    return dict(workload="kuramoto_em_parameter_table", rows=orbits, cols=len(dists), device_ms=ms,
                device_values_per_s=values / (ms * 1e-3), host_numpy_ms=host_ms,
                host_values_per_s=values / (host_ms * 1e-3), bytes_written=values * 8,
                device_GBps=values * 8 / (ms * 1e-3) / 1e9, finite=bool(np.isfinite(host).all()))


def batch_bench(solver, reps=1):
    """Brusselator chains of 16/64/256 cells (1:1:1), 32768 orbits, 500..5000
    steps each, implicit Euler (3) or midpoint (4) with at most 5 Newton
    updates per step, all in one persistent launch fed by a work queue."""
    import numpy as np
    import torch
    from sodecl_amd import configs, sodecl_interface as si
    cfg = configs.brusselator_batch(solver=solver)
    dev = torch.device("cuda:0")
    b = si.Batch(cfg["system"], solver, cfg["dt"], cfg["ksteps"], cfg["npar"], cfg["class_neq"], 0,
                 {"newton_max_iter": cfg["newton_max_iter"]})
    b.setup(cfg["cls"], cfg["steps"])
    y0 = torch.tensor(cfg["y0"], device=dev)
    p = torch.tensor(cfg["params"], device=dev)
    out = torch.empty((b.output_size(),), dtype=torch.float64, device=dev)
    nu = torch.zeros(cfg["orbits"], dtype=torch.int64, device=dev)
    stream = torch.cuda.Stream()
    b.run_device(y0.data_ptr(), p.data_ptr(), out.data_ptr(), nu.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        b.run_device(y0.data_ptr(), p.data_ptr(), out.data_ptr(), nu.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    kms = sum(ms) / len(ms)
    updates = nu.cpu().numpy()
    neq = np.asarray(cfg["class_neq"])[cfg["cls"]]
    flops = configs.brusselator_flops(neq, cfg["steps"], updates, solver)
    total = int(np.sum(cfg["steps"]))
    return dict(workload=f"brusselator_batch_{'ie' if solver == 3 else 'im'}", orbits=cfg["orbits"],
                classes_neq=cfg["class_neq"], steps_total=total, kernel_ms=kms, value=total / (kms * 1e-3),
                unit="orbit-steps/s", newton_updates=int(updates.sum()),
                state_updates=int(np.sum(neq * updates)), state_steps=int(np.sum(neq * cfg["steps"])),
                mean_updates_per_step=float(updates.sum() / total), achieved_tflops=flops / (kms * 1e-3) / 1e12,
                frac_fp64_peak=flops / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                finite=bool(torch.isfinite(out).all()))


def host_return_bench(cfg, reps=3):
    """sodecl_run of the workload with the whole trajectory returned to a
    fresh host array each run (the reference's contract and the Python
    wrapper's path): the integration streams chunk by chunk through pinned
    memory into the caller's array while the next chunk integrates.  Wall
    time of the run, PCIe and host copies included."""
    import numpy as np
    from sodecl_amd import configs, sodecl_interface as si
    steps = si.num_launches(cfg["tspan"], cfg["dt"], cfg["ksteps"]) * cfg["ksteps"]
    s = configs.session_for(cfg)
    s.setup(cfg["y0"], cfg["params"], None, cfg["counters"])
    out = s.run()
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = s.run()
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    return dict(workload=cfg["name"], path="sodecl_run, trajectory returned to a new host array",
                ms=wall * 1e3, value=cfg["orbits"] * steps / wall, unit="orbit-steps/s",
                trajectory_bytes=int(out.nbytes), trajectory_GBps=out.nbytes / wall / 1e9)


def host_share_bench(cfg, reps=3):
    """sodecl_run of the workload on the GPU alone and with a host share:
    the last orbits integrated on the host cores by the product's lane
    integrator compiled for the host, while the GPU integrates the rest.
    Both runs without trajectory output (the share changes only which
    orbits the GPU integrates; the trajectory's copy to host memory would
    add the same PCIe transfer to both).  The share is the balance point of
    the two measured rates, host orbit-steps/s (from a warmed probe run with
    four orbits per host thread) over host + GPU, rounded down to whole
    orbits per host thread (an orbit is a whole run of steps)."""
    import numpy as np
    from sodecl_amd import configs, sodecl_interface as si
    orbits = cfg["orbits"]
    steps = si.num_launches(cfg["tspan"], cfg["dt"], cfg["ksteps"]) * cfg["ksteps"]
    threads = len(os.sched_getaffinity(0))

    def session(**opts):
        s = configs.session_for(cfg, output_type=0, **opts)
        s.setup(cfg["y0"], cfg["params"], None, cfg["counters"])
        return s

    def median_wall(s):
        s.run()
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            s.run()
            walls.append(time.perf_counter() - t0)
        return float(np.median(walls)), s.last_kernel_ms()

    gpu = session()
    t_gpu, k_gpu = median_wall(gpu)
    del gpu
    probe = session(host_share=min(0.5, 4.0 * threads / orbits))
    probe.run()
    probe.run()
    p_ms, p_n = probe.host_stats()
    del probe
    host_rate = p_n * steps / (p_ms * 1e-3)
    gpu_rate = orbits * steps / (k_gpu * 1e-3)
    share = host_rate / (host_rate + gpu_rate)
    # whole orbits per thread: a thread's orbits run one after another
    per_thread = int(orbits * share / threads)
    res = dict(workload=cfg["name"], path="sodecl_run without trajectory output", host_threads=threads,
               host_orbit_steps_per_s=host_rate, gpu_kernel_orbit_steps_per_s=gpu_rate, balance_share=share,
               gpu_only_ms=t_gpu * 1e3, gpu_only_kernel_ms=k_gpu)
    if per_thread < 1:
        res["note"] = "the balance share is below one orbit per host thread: no host share"
        return res
    share = (per_thread * threads + 0.5) / orbits
    mix = session(host_share=share)
    t_mix, k_mix = median_wall(mix)
    h_ms, h_n = mix.host_stats()
    res.update(host_orbits=h_n, shared_ms=t_mix * 1e3, shared_kernel_ms=k_mix, host_ms=h_ms,
               gain=t_gpu / t_mix - 1.0)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this library is measured on one GPU; --gpus above 1 is not supported "
                                   "(orbits are independent: each GPU would integrate its own orbit range "
                                   "with no per-step communication)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from sodecl_amd import configs, sodecl_interface as si

    log(f"config {args.config}: building inputs")
    cfg = getattr(configs, args.config)()
    orbits = cfg["orbits"]
    opts = {"schedule": args.schedule}
    if args.group_size:
        opts["group_size"] = args.group_size
    if args.block_size:
        opts["block_size"] = args.block_size
    head, sess, traj = measure(cfg, args.steps, args.warmup, opts, args.config)
    value = head["value"]
    steps_per_run = head["steps_per_run"]
    sched = sess.launch_config()
    finite = head["finite"]

    plain = None
    if not args.no_plain:
        log("plain kernel (reference launch design) run")
        ps = configs.session_for(cfg, schedule="plain")
        ps.setup(cfg["y0"], cfg["params"], None, cfg["counters"])
        t0 = time.perf_counter()
        pres = ps.run()
        pwall = time.perf_counter() - t0
        plain = dict(value=orbits * steps_per_run / pwall, unit="orbit-steps/s", seconds=pwall,
                     design="lane per orbit, one launch per ksteps segment, blocking state read-back and "
                            "host transposition before each launch")
        got = traj.cpu().numpy().reshape(-1)
        same = bool(np.array_equal(np.nan_to_num(pres, nan=1e300), np.nan_to_num(got, nan=1e300)))
        plain["bitwise_equal_to_fast_path"] = same if sched[0] == 1 else None
        del pres, got

    if args.calibrate_copy_mb:
        log(f"calibration copy of {args.calibrate_copy_mb} MiB")
        calibration_copy(args.calibrate_copy_mb)

    adj = None
    if not args.no_adjoint and args.config == "oculomotor_rk4":
        log("forward+adjoint (Kuramoto N=64 RK4, 16384 orbits)")
        adj = fwd_adjoint_bench(max(1, min(args.steps, 3)))
        log("forward+adjoint with stored stage vectors (no recomputation)")
        stored = fwd_adjoint_bench(max(1, min(args.steps, 3)), store_stages=True)
        stored["cost_ratio_vs_forward"] = stored["fwd_adjoint_ms"] / adj["fwd_ms"]
        adj["stored_stages"] = stored

    batch = None
    if not args.no_batch and args.config == "oculomotor_rk4":
        log("batched implicit Brusselator (32768 orbits, mixed sizes and step counts)")
        batch = [batch_bench(3), batch_bench(4)]

    host_return = None
    if not args.no_host_return:
        log("sodecl_run with the trajectory returned to host memory")
        host_return = host_return_bench(cfg)
        if plain is not None:
            host_return["speedup_vs_plain"] = host_return["value"] / plain["value"]

    share = None
    if not args.no_host_share and args.config == "oculomotor_rk4":
        log("host share (sodecl_run on the GPU alone and with the host cores)")
        share = host_share_bench(cfg)

    samp = None
    if not args.no_sampling and args.config == "oculomotor_rk4":
        log("on-device parameter sampling (262144 x 65)")
        samp = sampling_bench()

    # the other explicit workloads, each with its own roofline and CPU sample
    legs = None
    if not args.no_legs and args.config == "oculomotor_rk4":
        legs = []
        for name, steps, warm, cpu_orbits in (("kuramoto_euler", 20, 3, 1024), ("kuramoto_em", 2, 1, 2048),
                                              ("oculomotor_rk4_plain", 5, 1, 0)):
            lcfg = getattr(configs, name)()
            leg, lsess, ltraj = measure(lcfg, steps, warm, {}, name)
            del lsess, ltraj
            leg["steps"] = steps
            leg["warmup"] = warm
            if not args.no_cpu_baseline and cpu_orbits:
                log(f"{name}: CPU baseline (reference kernels on host cores)")
                leg["cpu_baseline"] = cpu_baseline(lcfg, cpu_orbits)
                if leg["cpu_baseline"]:
                    leg["speedup_vs_cpu"] = leg["value"] / leg["cpu_baseline"]["value"]
            legs.append(leg)

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (reference kernels on host cores)")
        # about 5-15 s of wall time on 16 host threads
        sample = args.cpu_orbits or {"oculomotor_rk4": 65536, "kuramoto_euler": 1024, "kuramoto_em": 2048,
                                     "kuramoto_rk4": 512}[args.config]
        cpu = cpu_baseline(cfg, sample)

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "orbit-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded numpy PCG64 draws, sodecl_amd/configs.py)",
        "config": head["config"],
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "plain_kernel": plain,
        "fwd_adjoint": adj,
        "batch_implicit": batch,
        "param_sampling": samp,
        "host_return": host_return,
        "host_share": share,
        "workloads": legs,
        "speedup_vs_plain": None if plain is None else value / plain["value"],
        "speedup_vs_cpu": None if cpu is None else value / cpu["value"],
        "finite_fraction_final_state": finite,
        "tolerance": "1e-10 relative vs the CPU oracle (tests/test_gpu_parity.py)",
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
