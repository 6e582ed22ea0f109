"""Builds the in-tree shared libraries of the package.

* ``libsoapturbo_hip.so`` -- the HIP kernels and the C-ABI (gfx950 only);
* ``_workload.so``        -- the synthetic-input generator (host C++);
* ``libsoap_turbo_f.so``  -- the Fortran ``get_soap`` drop-in over the C-ABI
  (built when AMD flang is present).

``python -m soap_turbo_amd.build`` builds everything; the libraries are written
next to this file so that they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "lib", "llvm", "bin", "flang")
ARCH = "gfx950"

HIP_SOURCES = ["kernels_f64.hip", "adjoint.hip", "adjoint_f32.hip", "neighbours.hip", "gap.hip", "hessian.hip", "api.hip"]
CXX_SOURCES = ["basis.cpp"]
# packed-FP32 instructions are kept out of every HIP object
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
# The FP32 site kernel takes its next site's ticket with one lane and uses the result only at the
# next phase; the atomic optimizer's wave-wide form of that atomic would wait for it at once.
FILE_FLAGS = {"adjoint_f32.hip": ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]}


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build_hip(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        jobs.append(([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", path, "-o", obj], obj))
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(jobs) or 1) as pool:
        list(pool.map(lambda j: _run(j[0]), jobs))
    objs = [obj for _, obj in jobs]
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(HERE, "libsoapturbo_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-L" + os.path.join(ROCM, "lib"), "-ldl",
          "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_workload():
    out = os.path.join(HERE, "_workload.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", os.path.join(CSRC, "workload.cpp"), "-o",
          out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_fortran():
    """Fortran drop-in modules over the C-ABI (libsoap_turbo_f.so + .mod files in _build/fortran)."""
    if not os.path.exists(FLANG) or not os.path.isdir(FSRC):
        return None
    moddir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(moddir, exist_ok=True)
    srcs = [os.path.join(FSRC, f) for f in ("soap_turbo_c_api.f90", "soap_turbo_radial.f90",
                                            "soap_turbo_compress.f90", "soap_turbo_desc.f90", "soap_turbo_flat.f90")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return None
    out = os.path.join(HERE, "libsoap_turbo_f.so")
    _run([FLANG, "-O2", "-fPIC", "-shared", "-module-dir", moddir, *srcs, "-L" + HERE, "-lsoapturbo_hip",
          "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"], cwd=moddir)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    libs = [build_workload(), build_hip()]
    f = build_fortran()
    if f:
        libs.append(f)
    return libs


if __name__ == "__main__":
    for lib in build_all():
        print(lib)
