"""sodecl_amd: MI355X-native SODECL.

Integrates one user-supplied ODE/SDE system (OpenCL-C style source) for many
parameter sets ("orbits") at once with hand-written gfx950 integrators that
hipRTC compiles together with the system source.

Entry points:
  sodecl_amd.sodecl.sodecl(...)      the reference's Python wrapper contract
  sodecl_amd.sodecl_interface        the reference's pybind11 module
                                     (sodeclcall) plus Session
  sodecl_amd.torch_ops               device-resident runs on PyTorch tensors
The native library is libsodecl_amd.so (C ABI: include/sodecl.h).  There is
no CPU fallback: importing fails loudly when the library is not built.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SYSTEMS_DIR = os.path.join(HERE, "systems")

# One HIP runtime per process: PyTorch ships its own ROCm runtime libraries
# and loads them by path.  Importing it first lets libsodecl_amd.so bind to
# that same runtime (same sonames) instead of loading a second one.
try:  # pragma: no cover - depends on the environment
    import torch  # noqa: F401
except ImportError:
    pass

try:
    from . import sodecl_interface  # noqa: F401
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "sodecl_amd native library is not built (run python -m sodecl_amd.build): " + str(e)) from e

from . import sodecl  # noqa: E402,F401


def system_path(name):
    """Path of a bundled system plugin (oculomotor.cl, kuramoto.cl, brusselator.cl)."""
    if not name.endswith(".cl"):
        name += ".cl"
    return os.path.join(SYSTEMS_DIR, name)


def library_path():
    return os.path.join(HERE, "libsodecl_amd.so")
