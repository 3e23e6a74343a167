"""Builds the native parts of dsm_amd in-tree.

  libdsm_amd.so   C ABI (include/dsm.h): host engine + HIP kernels for gfx950
  bin/CDSM        command line shim over libdsm_amd.so
  libdsm_gen.so   seeded workload generator (host only)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("DSM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions anywhere (there is no floating point on the
    # path, but the flag keeps the compiler from emitting them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: %s" % " ".join(cmd))
    return out


def build_engine():
    objs = []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    for src in ("dsm_kernels.hip", "dsm_cluster.hip", "dsm_intersect.hip", "dsm_engine.cpp", "dsm_cluster_plan.cpp",
                "dsm_file.cpp"):
        obj = os.path.join(objdir, src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"] if src in ("dsm_engine.cpp", "dsm_cluster_plan.cpp") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-I", os.path.join(ROOT, "include"), "-c",
                                          os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(HERE, "libdsm_amd.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", lib] + objs)
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "cdsm_main.cpp"), "-o", os.path.join(bindir, "CDSM"),
          "-L", HERE, "-ldsm_amd", "-Wl,-rpath,$ORIGIN/.."])
    return lib


def build_gen():
    lib = os.path.join(HERE, "libdsm_gen.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", lib, os.path.join(CSRC, "gen.cpp")])
    return lib


def build_all():
    build_gen()
    build_engine()


if __name__ == "__main__":
    build_all()
