"""Build libfmm3d_mi355x.so (HIP for gfx950) and, when a Fortran compiler is
present, the ISO_C_BINDING drop-in library.  Run as `python -m
fmmlib3d_amd.build` or through __graft_entry__.build().

Builds are in-tree (fmmlib3d_amd/lib/) so that the shared objects travel with
the source tree to the GPU host.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(LIBDIR, "libfmm3d_mi355x.so")
FLIB = os.path.join(LIBDIR, "libfmm3d_mi355x_f.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CXX = os.path.join(ROCM, "llvm", "bin", "clang++")
FC = os.path.join(ROCM, "bin", "amdflang")

HIP_SOURCES = ["runtime.hip", "tree_gpu.hip", "fmm3d_lap.hip", "fmm3d_helm.hip", "fmm3d_api.hip", "server.hip",
               "htria_near.hip"]
# per-source flags: the Helmholtz kernels without machine LICM.  Hoisted
# constants of the inlined FP64 math routines otherwise cost the triangle
# near field half its waves (htria_near.hip) and the Helmholtz M2P 12-17%
# of its time (fmm3d_helm.hip); the results are the same bit for bit.
HIP_EXTRA = {"htria_near.hip": ["-mllvm", "-disable-machine-licm"],
             "fmm3d_helm.hip": ["-mllvm", "-disable-machine-licm"]}
CPP_SOURCES = ["octree.cpp", "tables.cpp", "helm_tables.cpp", "tria_rules.cpp"]

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fopenmp",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-march=x86-64-v3"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    for s in CPP_SOURCES:
        o = os.path.join(OBJDIR, s + ".o")
        _run([CXX] + CXX_FLAGS + ["-I", CSRC, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(OBJDIR, s + ".o")
        _run([HIPCC] + HIP_FLAGS + HIP_EXTRA.get(s, []) + ["-I", CSRC, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


FDRIVER = os.path.join(LIBDIR, "lfmm3dpart_gpu_driver")


def build_fortran(verbose=False):
    """ISO_C_BINDING library exporting the reference's F77 entry names, and a
    Fortran test driver linked against it."""
    if not os.path.exists(FC):
        return None
    src = os.path.join(HERE, "fortran", "fmm3d_mi355x.f90")
    os.makedirs(OBJDIR, exist_ok=True)
    tmp = FLIB + ".tmp"
    _run([FC, "-O2", "-fPIC", "-shared", "-module-dir", OBJDIR, src, "-o", tmp,
          "-L", LIBDIR, "-lfmm3d_mi355x", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    drv = os.path.join(ROOT, "tests", "fortran", "lfmm3dpart_gpu_driver.f90")
    tmp = FDRIVER + ".tmp"
    _run([FC, "-O2", "-module-dir", OBJDIR, drv, "-o", tmp, "-L", LIBDIR, "-lfmm3d_mi355x_f",
          "-lfmm3d_mi355x", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FDRIVER)
    if verbose:
        print("built", FLIB, "and", FDRIVER)
    return FLIB


def build_oracle(verbose=False):
    """The CPU oracle (test infrastructure) and, when the reference sources
    are present, the reference library compiled from them."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir])
    if os.path.isdir("/root/reference/src") and os.path.exists(FC):
        _run(["make", "-s", "-j8", "-C", odir, "ref"])
    if verbose:
        print("built oracle")


def build_test_gateway(verbose=False):
    """The MATLAB gateway compiled against the mx API test double in
    tests/matlab (test infrastructure; needs the HIP library built)."""
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "tests", "matlab")])
    if verbose:
        print("built MATLAB gateway test library")


def main():
    build_hip(verbose=True)
    build_fortran(verbose=True)
    build_oracle(verbose=True)
    build_test_gateway(verbose=True)


if __name__ == "__main__":
    main()
