"""Build steps of the native parts (no cmake/ninja needed).

* liblbmib.so -- the HIP library, hipcc for gfx950, in this package directory.
* lbmib_driver -- the thin Fortran driver (amdflang) that calls liblbmib
  through iso_c_binding, next to it.

Flags: ``-ffp-contract=off`` keeps every FP64 expression rounding exactly as
written (the drop-in path is checked bit for bit against a CPU restatement);
``-Xclang -target-feature -Xclang -packed-fp32-ops`` keeps the compiler from
emitting packed FP32 arithmetic.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["lbmib_kernels.hip", "lbmib_api.hip", "lbmib_engine.hip"]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
               "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    return r.stdout


def build_hip(verbose: bool = False) -> str:
    out = os.path.join(HERE, "liblbmib.so")
    tmp = out + ".tmp"
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    log = _run([hipcc] + HIPCC_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in HIP_SOURCES])
    if verbose:
        print(log)
    os.replace(tmp, out)
    return out


def build_fortran(verbose: bool = False) -> str:
    """Compile the iso_c_binding module and the driver against liblbmib.so."""
    flang = os.path.join(ROCM, "bin", "amdflang")
    if not os.path.exists(flang):
        flang = shutil.which("amdflang") or shutil.which("flang")
    if not flang:
        raise RuntimeError("no Fortran compiler (amdflang) found")
    out = os.path.join(HERE, "lbmib_driver")
    bdir = os.path.join(HERE, "_fbuild")
    os.makedirs(bdir, exist_ok=True)
    srcs = [os.path.join(FSRC, "lbmib_mod.f90"), os.path.join(FSRC, "lbmib_restart.f90"),
            os.path.join(FSRC, "lbmib_driver.f90")]
    _run([flang, "-O2", "-module-dir", bdir, "-o", out + ".tmp"] + srcs +
         ["-L" + HERE, "-llbmib", "-Wl,-rpath,$ORIGIN"], cwd=bdir)
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liblbmib_oracle.so")


if __name__ == "__main__":
    print(build_hip())
    print(build_oracle())
    try:
        print(build_fortran())
    except Exception as e:  # the Fortran driver is optional for the Python path
        print("fortran driver not built:", e)
