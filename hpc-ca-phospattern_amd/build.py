"""Build recipe for the calcium_sparks MI355X path.

Produces, under ``_build/`` next to this file:

* ``libcsparks.so``  -- HIP kernels + the cs_* C ABI (hipcc, gfx950):
  single domains (csparks.hip) and ensembles (cs_batch.hip)
* ``libcshost.so``   -- problem builder, Mersenne Twister, loop rules (gcc)
* ``libcsio.so``     -- HDF5 reader/writer for the reference file formats
* ``calcium_sparks`` -- the drop-in command-line program

Run ``python build.py`` (or ``__graft_entry__.build()``).  Everything is
compiled on every call; hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
INC = os.path.join(REPO, "include")
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
CONDA = os.environ.get("CS_HDF5_PREFIX", "/opt/conda")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return res.stdout


def _hipcc():
    path = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


DEVICE_SOURCES = ("csparks.hip", "cs_batch.hip", "cs_closure.hip")


def build_device():
    """Compile the HIP sources to objects side by side, then link them."""
    os.makedirs(OUT, exist_ok=True)
    out = os.path.join(OUT, "libcsparks.so")
    tmp = out + ".tmp"
    objs, procs = [], []
    for src in DEVICE_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [_hipcc(), *HIPCC_FLAGS, "-c", "-I", INC, "-I", CSRC, os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for cmd, p in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(log)
            raise RuntimeError("build step failed: " + " ".join(cmd))
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host():
    os.makedirs(OUT, exist_ok=True)
    out = os.path.join(OUT, "libcshost.so")
    srcs = [os.path.join(HOST, f) for f in ("cs_setup.c", "cs_mt.c", "cs_loop.c")]
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-I", INC, *srcs, "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def _hdf5_flags():
    """Link libhdf5 by path with a RUNPATH that only covers libcsio's own
    dependencies: a -L/-rpath on the whole prefix would also hand its older
    libstdc++ to the HIP runtime."""
    inc = os.path.join(CONDA, "include")
    lib = os.path.join(CONDA, "lib")
    so = os.path.join(lib, "libhdf5.so")
    if not os.path.exists(os.path.join(inc, "hdf5.h")) or not os.path.exists(so):
        raise RuntimeError(f"HDF5 not found under {CONDA}; set CS_HDF5_PREFIX")
    return ["-I", inc], [so, "-Wl,--enable-new-dtags", f"-Wl,-rpath,{lib}"]


def build_io():
    os.makedirs(OUT, exist_ok=True)
    out = os.path.join(OUT, "libcsio.so")
    cflags, ldflags = _hdf5_flags()
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-I", INC, *cflags,
          os.path.join(HOST, "cs_io.c"), "-o", out + ".tmp", *ldflags])
    os.replace(out + ".tmp", out)
    return out


def build_cli():
    """The calcium_sparks program, linked against the three libraries."""
    os.makedirs(OUT, exist_ok=True)
    out = os.path.join(OUT, "calcium_sparks")
    cflags, _ = _hdf5_flags()
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", INC, *cflags, os.path.join(HOST, "calcium_sparks.c"),
          "-o", out + ".tmp", "-L", OUT, "-lcshost", "-lcsio", "-lcsparks", "-Wl,--enable-new-dtags",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    """The CPU checker (oracle/) and, when the reference tree is present,
    the checker libraries compiled from its self-contained files."""
    odir = os.path.join(REPO, "oracle")
    _run(["make", "-s", "-C", odir, "all"])
    if os.path.isdir("/root/reference/fvm_mpi/src"):
        _run(["make", "-s", "-C", odir, "ref"])


def build_all():
    build_host()
    build_device()
    build_io()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built into", OUT)
