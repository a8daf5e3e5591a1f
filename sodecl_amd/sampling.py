"""On-device draws of per-orbit parameters and initial states.

    from sodecl_amd.sampling import sample
    params = sample([("uniform", 0, 5)] + [("normal", 1.0, 0.1)] * 32, orbits=262144, seed=7)

Column c of the [orbits][len(dists)] float64 result comes from dists[c]:
("constant", a), ("uniform", a, b), ("log_uniform", a, b) or
("normal", mean, deviation).  Entry (orbit, column) is drawn from one
Threefry-2x64-20 block with counter {orbit, column} and key
{seed, 0x5EEDD15C0FFEE5ED}, so draws do not depend on the launch shape and
any orbit range can be redrawn alone.  The arithmetic runs in the library's
gfx950 sampling kernel (sodecl_sample_device); PyTorch supplies the memory
and the stream.
"""
import torch

from . import sodecl_interface

KINDS = {"constant": 0, "uniform": 1, "log_uniform": 2, "normal": 3}


def table(dists):
    """(kind, a, b) triples of the C ABI from the readable form."""
    out = []
    for d in dists:
        kind = KINDS[d[0]]
        a = float(d[1])
        b = float(d[2]) if len(d) > 2 else 0.0
        out.append((kind, a, b))
    return out


def sample(dists, orbits, seed, device=0, out=None):
    dev = torch.device("cuda", device)
    t = table(dists)
    if out is None:
        out = torch.empty((orbits, len(t)), dtype=torch.float64, device=dev)
    elif out.dtype != torch.float64 or tuple(out.shape) != (orbits, len(t)) or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float64 tensor of shape {(orbits, len(t))}")
    stream = torch.cuda.current_stream(dev).cuda_stream
    sodecl_interface.sample_device(t, orbits, int(seed), out.data_ptr(), stream)
    return out
