"""Build libqasmhd_amd.so (HIP, gfx950) in-tree: qasmhd_amd/_lib/.

    python -m qasmhd_amd.build        # or __graft_entry__.build()

Every .hip file under csrc/ is compiled with hipcc for gfx950 and linked
into one shared library exporting the C ABI of include/qasmhd.h.  Flags:
  -ffp-contract=off      no FMA contraction: the FP64 path keeps the
                         reference's rounding sequence (bit-identical results);
                         qas_mixed.hip and qas_fma64.hip (precision 1, 2)
                         are built with contraction, and the discrete adjoint
                         (qas_adjoint2d.hip) with contraction within each
                         source expression (-ffp-contract=on) and FP64
                         division as reciprocal + Newton steps
                         (-freciprocal-math -fapprox-func)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions are generated
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
OBJDIR = os.path.join(HERE, "_lib", "obj")
LIBNAME = "libqasmhd_amd.so"
ARCH = os.environ.get("QAS_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-Wno-unused-command-line-argument", "-Wno-unused-result"]
# per-source flags: every unit is built without FMA contraction except the
# FP32 sweeps of the mixed path, which also use the fast FP32 divide/sqrt,
# the contracted FP64 sweeps (IEEE FP64 divide and sqrt), and the discrete
# adjoint.  The adjoint is checked against finite differences, not bit for
# bit; "on" fuses a*b+c only inside one source expression, so every kernel
# that inlines the same code (single-domain and batched) rounds alike.  Its
# divisions are a reciprocal estimate refined by Newton steps and a
# corrected quotient, without the operand scaling and special-case fix-up
# of the IEEE sequence (its operands are densities, pressures, wave speeds
# and smoothness sums, far from the exponent range's ends)
EXTRA = {"qas_mixed.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
         "qas_fma64.hip": ["-ffp-contract=fast"],
         "qas_adjoint2d.hip": ["-ffp-contract=on", "-freciprocal-math", "-fapprox-func"]}


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, src[:-4] + ".o")
    flags = list(COMMON_FLAGS) + EXTRA.get(src, ["-ffp-contract=off"])
    cmd = [_hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library.  Always
    recompiles (this is the build check)."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = lib_path()
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, out)
    if verbose:
        print(out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
