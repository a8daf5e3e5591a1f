"""Device Threefry-2x64-20 against known answers, all 128 bits, through the
C ABI's diagnostic entry point (sodecl_threefry_device), which runs the same
device functions as the SDE integrators and the sampler.

Vectors:
  * Random123's published threefry2x64 (20 rounds) known-answer vectors
    (zero, all-ones and pi-digit inputs);
  * the reference SDE caller's key {0, 0xf00dcafe} with counter {orbit,
    counter} (src/kernels/stochastic_solver_caller.cl:31-52), output words
    computed by compiling the reference's vendored Random123 header
    (Random123/threefry.h:178-293) as C, together with the open-interval
    doubles of the first block (u01.h:129-131);
  * 200000 random (counter, key) pairs against the oracle's Threefry, which
    the published vectors pin on the CPU (tests/test_oracle.py).
"""
import ctypes

import numpy as np
import pytest
import torch

import sodecl_amd
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ONES = 0xFFFFFFFFFFFFFFFF
REF_KEY = (0, 0xF00DCAFE)
# (ctr0, ctr1, key0, key1) -> (word0, word1)
KATS = [
    ((0, 0, 0, 0), (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)),
    ((ONES, ONES, ONES, ONES), (0xE02CB7C4D95D277A, 0xD06633D0893B8B68)),
    ((0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0, 0x082EFA98EC4E6C89),
     (0x263C7D30BB0F0AF1, 0x56BE8361D3311526)),
    ((0, 0) + REF_KEY, (0x03ED3FEB0846F822, 0xE6E400C11521CD65)),
    ((0, 1) + REF_KEY, (0x073882EA45742B79, 0x8B11D13AB0B7FBD6)),
    ((1, 0) + REF_KEY, (0x2350FE2D7FE3C86D, 0x89D9E9C8C90C4350)),
    ((1, 1) + REF_KEY, (0xC2C9BCB296ABC516, 0xCA58BCD111A5AE93)),
]


def _lib():
    lib = ctypes.CDLL(sodecl_amd.library_path())
    lib.sodecl_threefry_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.sodecl_threefry_device.restype = ctypes.c_int
    return lib


def device_blocks(inputs):
    """inputs: uint64 array [n][4] -> uint64 array [n][4] from the GPU."""
    inp = np.ascontiguousarray(inputs, dtype=np.uint64)
    n = inp.shape[0]
    d_in = torch.from_numpy(inp.view(np.int64)).to("cuda:0")
    d_out = torch.full((n, 4), -1, dtype=torch.int64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    assert _lib().sodecl_threefry_device(d_in.data_ptr(), d_out.data_ptr(), n, stream) == 1
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(np.uint64)


def test_published_and_reference_key_vectors_bit_exact():
    inputs = np.array([k[0] for k in KATS], dtype=np.uint64)
    out = device_blocks(inputs)
    for (inp, want), got in zip(KATS, out):
        assert (int(got[0]), int(got[1])) == want, f"input {[hex(v) for v in inp]}: got {hex(got[0])} {hex(got[1])}"


def test_reference_key_open_interval_doubles():
    out = device_blocks(np.array([[0, 0, REF_KEY[0], REF_KEY[1]]], dtype=np.uint64))
    u = out[0, 2:4].copy().view(np.float64)
    # (1 + (w >> 11)) * (2^53 - 1) * 2^-106 of the two words above
    assert u[0] == 0.015338892823162807
    assert u[1] == 0.90191654886178363
    for w, v in zip(KATS[3][1], u):
        assert v == float(1 + (w >> 11)) * float.fromhex("0x1.fffffffffffffp-54")


def test_random_inputs_match_oracle():
    rng = np.random.default_rng(123)
    n = 200000
    inputs = rng.integers(0, 2 ** 63, (n, 4), dtype=np.int64).view(np.uint64)
    inputs[:, :2] ^= rng.integers(0, 2, (n, 2), dtype=np.int64).view(np.uint64) << np.uint64(63)
    inputs[: n // 4, 2:] = np.array(REF_KEY, dtype=np.uint64)   # the SDE key, many counters
    out = device_blocks(inputs)
    o = Oracle("kuramoto")
    check = np.linspace(0, n - 1, 4000).astype(int)
    for i in check:
        c = inputs[i]
        w = o.threefry((int(c[0]), int(c[1])), (int(c[2]), int(c[3])))
        assert (int(out[i, 0]), int(out[i, 1])) == w, i
        assert out[i, 2:4].copy().view(np.float64)[0] == o.u01(w[0])
