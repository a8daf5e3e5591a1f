"""Builds the native parts of sodecl_amd in-tree.

  sodecl_amd/libsodecl_amd.so          C ABI (include/sodecl.h): host manager,
                                       hipRTC JIT, ahead-of-time gfx950 kernels
  sodecl_amd/sodecl_interface*.so      pybind11 module (reference entry point)
  sodecl_amd/_jit_cache/*.hsaco        warmed code objects of the bundled systems
  oracle/build, oracle/_ref            CPU oracle (test infrastructure)

Usage: python -m sodecl_amd.build [--no-oracle] [--no-warm]
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("SODECL_AMD_ARCH", "gfx950")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
# Every HIP compile: no packed FP32 instructions.
NO_PK_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
DEVICE_SOURCES = ["random.hip", "gauss.hip", "integrators.hip", "adjoint.hip"]
HOST_SOURCES = ["jit.cpp", "sodecl_capi.cpp", "sodeclmgr.cpp", "adjoint.cpp", "batch.cpp", "host_share.cpp"]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def _raw_string(var, names):
    parts = []
    for name in names:
        path = os.path.join(CSRC, "device", name)
        if os.path.exists(path):
            parts.append(f'#line 1 "{name}"\n' + open(path).read())
    text = "\n".join(parts)
    delim = "SODECL_EMBED"
    assert f"){delim}\"" not in text
    # keep long strings under compiler limits by concatenating chunks
    chunks = [text[i:i + 60000] for i in range(0, len(text), 60000)]
    body = "\n".join(f'R"{delim}({c}){delim}"' for c in chunks)
    return f"static const char* const {var} =\n{body};\n"


def embed_device_sources():
    """device/*.hip -> _embedded_device.inc (C++ raw strings): the prelude
    emitted before the system source and the integrators emitted after it."""
    out = os.path.join(CSRC, "_embedded_device.inc")
    new = (_raw_string("kPreludeSource", ["sin.hip", "exp.hip", "prelude.hip"]) + _raw_string("kIntegratorsSource", DEVICE_SOURCES) +
           _raw_string("kHostMathSource", ["sin.hip", "exp.hip"]) +
           _raw_string("kHostLaneSource", ["random.hip", "gauss.hip", "integrators.hip"]) +
           _raw_string("kBatchClassSource", ["batch_class.hip"]) +
           _raw_string("kBatchKernelSource", ["batch_kernel.hip"]))
    if not os.path.exists(out) or open(out).read() != new:
        with open(out, "w") as f:
            f.write(new)


def build_library():
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for s in srcs:
        o = os.path.join(HERE, "_build", os.path.basename(s) + ".o")
        run(["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result", "-D__HIP_PLATFORM_AMD__",
             f"-I{ROCM}/include", "-c", s, "-o", o])
        objs.append(o)
    hip_obj = os.path.join(HERE, "_build", "static_kernels.o")
    run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", *NO_PK_FP32, "-c",
         os.path.join(CSRC, "static_kernels.hip"), "-o", hip_obj])
    objs.append(hip_obj)
    lib = os.path.join(HERE, "libsodecl_amd.so")
    # linked under a temporary name and renamed: a reader never sees a
    # partly written library
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", lib + ".tmp", f"-L{ROCM}/lib", "-lhiprtc",
         "-lamdhip64", "-ldl", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(lib + ".tmp", lib)
    return lib


def build_pybind():
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "sodecl_interface" + suffix)
    py_inc = sysconfig.get_paths()["include"]
    run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", f"-I{pybind11.get_include()}", f"-I{py_inc}",
         os.path.join(CSRC, "pybind_module.cpp"), f"-L{HERE}", "-lsodecl_amd", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8", "all"])


def warm_cache():
    """Compiles the bundled benchmark configurations into _jit_cache so the
    first run on a GPU skips hipRTC."""
    sys.path.insert(0, ROOT)
    from sodecl_amd import configs
    configs.warm_jit_cache(arch=ARCH)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    embed_device_sources()
    build_library()
    build_pybind()
    if "--no-oracle" not in argv:
        build_oracle()
    if "--no-warm" not in argv:
        warm_cache()


if __name__ == "__main__":
    main()
