"""Builds libgpvmc_hip.so in-tree for gfx950 with hipcc (no JIT cache, no torch)."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib", "libgpvmc_hip.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]

SOURCES = ("vmc_kernels.hip", "exc_kernels.hip", "exc_reg_kernels.hip", "detinv_kernels.hip", "gpvmc_api.cpp")


def sources():
    return [os.path.join(CSRC, f) for f in SOURCES]


def build(verbose=False):
    """Compiles every source to an object in parallel, then links the shared library."""
    objdir = os.path.join(os.path.dirname(LIB), "obj")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc] + HIP_FLAGS + ["-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed building libgpvmc_hip.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libgpvmc_hip.so")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST_LIB = os.path.join(PKG, "lib", "libgpvmc_host.so")
HOST_EXE = os.path.join(PKG, "lib", "vmc_host")
CONCAT_EXE = os.path.join(PKG, "lib", "concatenate")


# HDF5 (C library and its high-level API) of the image, for the reference's output files
HDF5 = os.environ.get("GPVMC_HDF5", "/opt/conda")


def build_host(verbose=False):
    """The C++ host classes (include/gpvmc_host.hpp) as libgpvmc_host.so over
    libgpvmc_hip.so and HDF5, the vmc_host driver and the concatenate tool; plain C++ (no
    device code)."""
    cxx = os.environ.get("CXX", "g++")
    libdir = os.path.dirname(LIB)
    host = os.path.join(PKG, "host")
    common = ["-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include")]
    h5lib = os.path.join(HDF5, "lib")
    cmds = [
        [cxx] + common + ["-I" + os.path.join(HDF5, "include"), "-fPIC", "-shared", "-o", HOST_LIB + ".tmp",
                          os.path.join(host, "gpvmc_host.cpp"), os.path.join(host, "h5merge.cpp"),
                          "-L" + libdir, "-lgpvmc_hip",
                          os.path.join(h5lib, "libhdf5_hl.so"), os.path.join(h5lib, "libhdf5.so"),
                          # the system library directory before HDF5's: the system libstdc++ (the
                          # one ROCm needs), not the older one next to HDF5
                          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu", "-Wl,-rpath," + h5lib],
        [cxx] + common + ["-o", HOST_EXE + ".tmp", os.path.join(host, "vmc_host.cpp"), "-L" + libdir,
                          "-lgpvmc_host", "-lgpvmc_hip", "-Wl,-rpath,$ORIGIN"],
        [cxx] + common + ["-o", CONCAT_EXE + ".tmp", os.path.join(host, "concatenate.cpp"), "-L" + libdir,
                          "-lgpvmc_host", "-lgpvmc_hip", "-Wl,-rpath,$ORIGIN"],
    ]
    for cmd, out in zip(cmds, (HOST_LIB, HOST_EXE, CONCAT_EXE)):
        if verbose:
            print(" ".join(cmd), flush=True)
        if out == HOST_EXE:
            os.replace(HOST_LIB + ".tmp", HOST_LIB)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("building %s failed" % os.path.basename(out))
    os.replace(HOST_EXE + ".tmp", HOST_EXE)
    os.replace(CONCAT_EXE + ".tmp", CONCAT_EXE)


if __name__ == "__main__":
    print(build(verbose=True))
