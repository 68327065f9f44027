"""Builds libmpsft_hip.so in-tree with hipcc for gfx950, and beside it
libmpsft_torch.so, the torch operator of the differentiable recovery (host
C++ against libtorch, calling libmpsft_hip.so through its C-ABI).

Run as ``python -m mpsft_amd.build`` (or through ``__graft_entry__.build``).
Both libraries are always rebuilt from source.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpsft_hip.so")
TORCH_LIB = os.path.join(HERE, "libmpsft_torch.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
SOURCES = ["kernels.hip", "harness.hip", "engine.cpp", "batch.cpp", "capi.cpp"]
# harness.hip restates the reference's signal generator, whose host
# normal_distribution must round as the reference's g++ build (no FMA).
SOURCE_FLAGS = {"harness.hip": ["-ffp-contract=off"]}
ARCH = os.environ.get("MPSFT_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions stay off (they are produced by hipcc -O3 from
# ordinary float code otherwise).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def build(verbose=False):
    """Compile every source and link the shared library.  Returns its path."""
    objs = []
    tmp = tempfile.mkdtemp(prefix="mpsft_build_")
    try:
        for src in SOURCES:
            obj = os.path.join(tmp, src + ".o")
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            cmd = [hipcc()] + HIP_FLAGS + SOURCE_FLAGS.get(src, []) + lang + \
                ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            err = "\n".join(l for l in r.stderr.splitlines()
                            if "packed-fp32-ops" not in l and l.strip())
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n{err}")
            if err and verbose:
                print(err, file=sys.stderr)
            objs.append(obj)
        out_tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC",
               "-Wl,-soname,libmpsft_hip.so", "-o", out_tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr}")
        os.replace(out_tmp, LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


def build_torch(verbose=False):
    """Compile csrc/torch_recover.cpp with g++ against the installed torch
    (its headers, C++ ABI and libraries) and link it to libmpsft_hip.so,
    found at run time beside it.  Returns its path."""
    import torch
    import torch.utils.cpp_extension as ext
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-Wall", "-Wno-unused-variable"]
    cmd += [f"-I{p}" for p in ext.include_paths()] + [f"-I{INCLUDE}"]
    cmd += [os.path.join(CSRC, "torch_recover.cpp")]
    for p in ext.library_paths():
        cmd += [f"-L{p}", f"-Wl,-rpath,{p}"]
    out_tmp = TORCH_LIB + ".tmp"
    cmd += ["-ltorch", "-ltorch_cpu", "-lc10", f"-L{HERE}", "-l:libmpsft_hip.so",
            "-Wl,-rpath,$ORIGIN", "-o", out_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on torch_recover.cpp:\n{r.stderr[-4000:]}")
    os.replace(out_tmp, TORCH_LIB)
    return TORCH_LIB


CLIENT_SRC = os.path.join(os.path.dirname(HERE), "tests", "capi_client.cpp")
CLIENT = os.path.join(os.path.dirname(HERE), "tests", "capi_client")


def build_capi_client(verbose=False):
    """Compile tests/capi_client.cpp, a C++11 caller of the C-ABI, with g++
    against include/mpsft.h only (no HIP headers) and link it to
    libmpsft_hip.so, as a reference-side C++ caller would.  The GPU tests
    run it.  Returns its path."""
    out_tmp = CLIENT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++11", "-Wall", f"-I{INCLUDE}", CLIENT_SRC, f"-L{HERE}",
           "-l:libmpsft_hip.so", "-Wl,-rpath,$ORIGIN/../mpsft_amd", "-o", out_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on capi_client.cpp:\n{r.stderr[-4000:]}")
    os.replace(out_tmp, CLIENT)
    return CLIENT


def build_gather_probe(verbose=False):
    """Compile tests/gather_probe.hip (random-gather load forms, a hardware
    measurement behind the gather-fold's line-size figures) with hipcc for
    gfx950.  Returns its path."""
    tests = os.path.join(os.path.dirname(HERE), "tests")
    src = os.path.join(tests, "gather_probe.hip")
    out = os.path.join(tests, "gather_probe")
    cmd = [hipcc()] + HIP_FLAGS + [src, "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on gather_probe.hip:\n{r.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_torch(verbose=True))
    print(build_capi_client(verbose=True))
    print(build_gather_probe(verbose=True))
