"""Shared comparison rules for the parity tests.

Tolerance for FP64 trajectories: 1e-10 relative.  An element passes when
|gpu - oracle| <= 1e-10 * max(|oracle|, scale), scale being the largest
finite |oracle| of that equation of that orbit over the compared snapshots
(so a component crossing zero is judged against its own magnitude along the
orbit).  Non-finite values must be non-finite in both.
"""
import numpy as np

RTOL = 1e-10


def rel_err(gpu, ref):
    """Elementwise error relative to max(|ref|, per-equation scale);
    arrays [snap][eq][orbit]."""
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    finite = np.isfinite(ref)
    scale = np.max(np.where(finite, np.abs(ref), 0.0), axis=0, keepdims=True)
    scale = np.maximum(scale, 1e-300)
    with np.errstate(invalid="ignore"):
        err = np.abs(gpu - ref) / np.maximum(np.abs(ref), scale)
    return np.where(finite & np.isfinite(gpu), err, np.where(finite == np.isfinite(gpu), 0.0, np.inf))


def assert_traj_close(gpu, ref, rtol=RTOL, what=""):
    err = rel_err(gpu, ref)
    worst = float(np.max(err)) if err.size else 0.0
    assert worst <= rtol, f"{what}: max relative error {worst:.3e} > {rtol:.0e}"
    return worst


def orbit_errors(gpu, ref):
    """Worst relative error per orbit (inf where finiteness differs)."""
    return rel_err(gpu, ref).max(axis=(0, 1))
