"""Build recipes for the parity oracle (TEST INFRASTRUCTURE).

Two things are built here, both only for tests/, smoke() and bench.py's
cpu_baseline leg:

* ``oracle/_build/libqas_oracle.so`` -- the C restatement in qas_oracle.c,
  compiled with the reference's flags (``-O2 -fopenmp``, Makefile.inc:2) plus
  ``-ffp-contract=off`` so no FMA is ever formed.
* ``oracle/_ref/`` -- the reference's OWN sources compiled from where they lie
  under /root/reference (common/*.c, the 1D/2D serial drivers) and linked with
  the small harnesses in oracle/ref/, which subclass the drivers to run an
  exact number of steps on a state read from disk.  Grid sizes and solver
  switches are the reference's compile-time macros (mymacros.hpp); the
  harness build defines ``_MYMACROS_`` so the driver's own mymacros.hpp is
  skipped by its include guard and the values below are used instead.
  Nothing from the reference is copied; only the build outputs land in
  oracle/_ref/.  The 3D driver needs mpi.h, which this image lacks, so it is
  not built; the 3D oracle is pinned by tests/test_oracle_props.py.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("QAS_REFERENCE", "/root/reference")
BUILD = os.path.join(HERE, "_build")
REFOUT = os.path.join(HERE, "_ref")
REF_CXX = ["g++", "-O2", "-fopenmp"]  # Makefile.inc:1-2
REF_COMMON = ["common_func", "common_interp", "mhd_eigen", "mhd_flux", "mhd_func"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def oracle_lib_path() -> str:
    return os.path.join(BUILD, "libqas_oracle.so")


def build_oracle(force: bool = True) -> str:
    """Compile the C restatement (always recompiles when force)."""
    os.makedirs(BUILD, exist_ok=True)
    out = oracle_lib_path()
    src = os.path.join(HERE, "qas_oracle.c")
    if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        tmp = out + ".tmp%d" % os.getpid()
        _run(["gcc", "-std=gnu99", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
              "-o", tmp, src, "-lm"])
        os.replace(tmp, out)
    return out


def reference_present() -> bool:
    return os.path.isdir(os.path.join(REF, "common"))


def _ref_objects(force: bool = False):
    objdir = os.path.join(REFOUT, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for name in REF_COMMON:
        src = os.path.join(REF, "common", name + ".c")
        obj = os.path.join(objdir, name + ".o")
        if force or not os.path.exists(obj):
            _run(REF_CXX + ["-x", "c++", "-c", src, "-o", obj])
        objs.append(obj)
    return objs


def build_ref_unit(force: bool = False) -> str:
    out = os.path.join(REFOUT, "unit")
    if force or not os.path.exists(out):
        objs = _ref_objects(force)
        _run(REF_CXX + ["-I", os.path.join(REF, "common"), os.path.join(HERE, "ref", "unit.cpp")] + objs
             + ["-lm", "-o", out])
    return out


def build_ref_merge(force: bool = False) -> str:
    """The reference's merge tool of the 2D MPI drivers (2D/MPI/common/merge.c:
    per-rank outdat_%05d_%05d.dat -> merge_outdat_%05d.dat without ghosts),
    compiled from its one source file."""
    os.makedirs(REFOUT, exist_ok=True)
    out = os.path.join(REFOUT, "merge2d")
    if force or not os.path.exists(out):
        _run(["gcc", "-O2", os.path.join(REF, "2D", "MPI", "common", "merge.c"), "-o", out])
    return out


def _cfl_tag(cfl: float) -> str:
    return ("%g" % cfl).replace(".", "p").replace("-", "m")


def ref_binary_name(ndim: int, nx: int, ny: int, rmn: int, odr: int, r_k: int, ctw: int, cfl: float,
                    gravity: bool = False) -> str:
    if ndim == 1:
        return "ref1d_%d_r%do%dk%d_c%s" % (nx, rmn, odr, r_k, _cfl_tag(cfl))
    return "ref2d%s_%dx%d_r%do%dk%dw%d_c%s" % ("g" if gravity else "", nx, ny, rmn, odr, r_k, ctw, _cfl_tag(cfl))


def build_ref(ndim: int, xmesh: int, ymesh: int = 1, rmn: int = 1, odr: int = 4, r_k: int = 3, ctw: int = 1,
              cfl: float = 0.4, force: bool = False, gravity: bool = False) -> str:
    """Build (once) the reference 1D or 2D driver for one macro set.

    xmesh/ymesh are the reference's XMESH/YMESH (interior cells); gravity
    builds the 2D driver with gravity (2D/SERIAL/RTI) instead of OTvortex."""
    os.makedirs(REFOUT, exist_ok=True)
    name = ref_binary_name(ndim, xmesh, ymesh, rmn, odr, r_k, ctw, cfl, gravity)
    out = os.path.join(REFOUT, name)
    if os.path.exists(out) and not force:
        return out
    objs = _ref_objects()
    macros = ["-D_MYMACROS_", "-DXMESH=(%d)" % xmesh, "-DN_OUT=(1)", "-DDTREC=(1.0)", "-DCFL=(%r)" % cfl,
              "-DRMN=(%d)" % rmn, "-DODR=(%d)" % odr, "-DR_K=(%d)" % r_k]
    if ndim == 1:
        ddir = os.path.join(REF, "1D", "SERIAL", "shock")
        srcs = [os.path.join(HERE, "ref", "harness1d.cpp"), os.path.join(ddir, "mhd1d_class.cpp"),
                os.path.join(ddir, "mhd1d_solve.cpp")]
    else:
        ddir = os.path.join(REF, "2D", "SERIAL", "RTI" if gravity else "OTvortex")
        macros += ["-DYMESH=(%d)" % ymesh, "-DCTW=(%d)" % ctw]
        if gravity:
            macros += ["-DRANDOM=(0)", "-DQAS_HARNESS_GRAVITY"]
        srcs = [os.path.join(HERE, "ref", "harness2d.cpp"), os.path.join(ddir, "mhd2d_class.cpp"),
                os.path.join(ddir, "mhd2d_solve.cpp")]
    tmp = out + ".tmp%d" % os.getpid()
    _run(REF_CXX + macros + ["-I", os.path.join(REF, "common"), "-I", ddir, "-I", os.path.join(HERE, "ref")]
         + srcs + objs + ["-lm", "-o", tmp])
    os.replace(tmp, out)
    return out


def ref_dsptv_name(nx: int, ny: int, rmn: int, odr: int, r_k: int, ctw: int) -> str:
    return "refdsp2d_%dx%d_r%do%dk%dw%d" % (nx, ny, rmn, odr, r_k, ctw)


def build_ref_dsptv(xmesh: int, ymesh: int, rmn: int = 1, odr: int = 4, r_k: int = 3, ctw: int = 1,
                    force: bool = False) -> str:
    """Build (once) the reference's dissipative 2D driver (2D/SERIAL/MRX:
    DMHD2D::dsptv with MHD2D::ideal/bound) with the dsptv harness."""
    os.makedirs(REFOUT, exist_ok=True)
    out = os.path.join(REFOUT, ref_dsptv_name(xmesh, ymesh, rmn, odr, r_k, ctw))
    if os.path.exists(out) and not force:
        return out
    objs = _ref_objects()
    ddir = os.path.join(REF, "2D", "SERIAL", "MRX")
    macros = ["-D_MYMACROS_", "-DXMESH=(%d)" % xmesh, "-DYMESH=(%d)" % ymesh, "-DN_OUT=(1)", "-DDTREC=(1.0)",
              "-DCFL=(0.4)", "-DRMN=(%d)" % rmn, "-DODR=(%d)" % odr, "-DR_K=(%d)" % r_k, "-DCTW=(%d)" % ctw,
              "-DRANDOM=(0)", "-DREV=(2e2)", "-DREM=(2e2)"]
    srcs = [os.path.join(HERE, "ref", "harness_dmhd2d.cpp")] + [
        os.path.join(ddir, f) for f in ("dmhd2d_class.cpp", "dmhd2d_solve.cpp", "dmhd2d_init_.cpp",
                                        "mhd2d_class.cpp", "mhd2d_solve.cpp")]
    tmp = out + ".tmp%d" % os.getpid()
    _run(REF_CXX + macros + ["-I", os.path.join(REF, "common"), "-I", ddir, "-I", os.path.join(HERE, "ref")]
         + srcs + objs + ["-lm", "-o", tmp])
    os.replace(tmp, out)
    return out


def run_ref_dsptv(binary: str, state, nu, eta, dx: float, dy: float, gam: float, dt: float, nsub: int = 1,
                  mode: int = 0, nsteps: int = 0, dcmax: float = 0.0, dnx=(0,) * 8, dny=(0,) * 8,
                  threads: int | None = None, timing: bool = False):
    """Run the dsptv harness; returns the final state (and the wall time of
    its loop in seconds when timing is set)."""
    import numpy as np
    state = np.ascontiguousarray(state, dtype=np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout, fnu, feta, fpar = (os.path.join(td, n) for n in ("in.bin", "out.bin", "nu.bin", "eta.bin",
                                                                      "p.txt"))
        state.tofile(fin)
        np.ascontiguousarray(nu, dtype=np.float64).tofile(fnu)
        np.ascontiguousarray(eta, dtype=np.float64).tofile(feta)
        with open(fpar, "w") as f:
            f.write("in %s\nout %s\nnu %s\neta %s\n" % (fin, fout, fnu, feta))
            f.write("dx %s\ndy %s\ngam %s\ndt %s\n" % (_hex(dx), _hex(dy), _hex(gam), _hex(dt)))
            f.write("nsub %d\nmode %d\nnsteps %d\ndcmax %s\n" % (nsub, mode, nsteps, _hex(dcmax)))
            f.write("dnx %s\ndny %s\n" % (" ".join(map(str, dnx)), " ".join(map(str, dny))))
        env = dict(os.environ)
        if threads is not None:
            env["OMP_NUM_THREADS"] = str(threads)
        r = subprocess.run([binary, fpar], capture_output=True, text=True, cwd=td, env=env)
        if r.returncode != 0:
            raise RuntimeError("reference dsptv harness failed (%d): %s %s" % (r.returncode, r.stdout, r.stderr))
        out = np.fromfile(fout, dtype=np.float64).reshape(state.shape)
    if timing:
        el = float("nan")
        for line in r.stdout.splitlines():
            if line.startswith("elapsed_s"):
                el = float(line.split()[1])
        return out, el
    return out


def _hex(x: float) -> str:
    return float(x).hex()


def run_ref(binary: str, state, dx: float, gam: float, nsteps: int, flg: int, dt: float = 0.0,
            dnx=(0,) * 8, dny=(0,) * 8, dy: float = 1.0, threads: int | None = None, timing: bool = False,
            phi=None):
    """Run a reference harness on `state` (numpy (8, ...) float64, ghosts
    included).  Returns (final_state, dt_history), plus the wall time of the
    step loop in seconds when timing is set."""
    import numpy as np
    state = np.ascontiguousarray(state, dtype=np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout, fdt, fpar = (os.path.join(td, n) for n in ("in.bin", "out.bin", "dt.bin", "p.txt"))
        state.tofile(fin)
        with open(fpar, "w") as f:
            f.write("in %s\nout %s\ndthist %s\n" % (fin, fout, fdt))
            if phi is not None:
                fphi = os.path.join(td, "phi.bin")
                np.ascontiguousarray(phi, dtype=np.float64).tofile(fphi)
                f.write("phi %s\n" % fphi)
            f.write("dx %s\ndy %s\ngam %s\n" % (_hex(dx), _hex(dy), _hex(gam)))
            f.write("nsteps %d\nflg %d\ndt %s\n" % (nsteps, flg, _hex(dt)))
            f.write("dnx %s\ndny %s\n" % (" ".join(map(str, dnx)), " ".join(map(str, dny))))
        env = dict(os.environ)
        if threads is not None:
            env["OMP_NUM_THREADS"] = str(threads)
        r = subprocess.run([binary, fpar], capture_output=True, text=True, env=env, cwd=td)
        if r.returncode != 0:
            raise RuntimeError("reference harness failed (%d): %s %s" % (r.returncode, r.stdout, r.stderr))
        out = np.fromfile(fout, dtype=np.float64).reshape(state.shape)
        hist = np.fromfile(fdt, dtype=np.float64) if os.path.exists(fdt) else np.zeros(0)
    if timing:
        el = float("nan")
        for line in r.stdout.splitlines():
            if line.startswith("elapsed_s"):
                el = float(line.split()[1])
        return out, hist, el
    return out, hist


def run_unit(func: str, sel: int, inp, gam: float = 5.0 / 3.0):
    """Evaluate a reference routine on a batch (see oracle/ref/unit.cpp)."""
    import numpy as np
    binary = build_ref_unit()
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[0]
    width = {"interp": 2, "fcen": 1, "df": 1, "lrstate": 14, "flux": 7, "cuct": 1}[func]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        inp.tofile(fin)
        _run([binary, func, str(sel), str(n), _hex(gam), fin, fout])
        return np.fromfile(fout, dtype=np.float64).reshape(n, width)


if __name__ == "__main__":
    print(build_oracle())
    if reference_present():
        print(build_ref_unit())
        print(build_ref(2, 16, 16))
        print(build_ref(1, 64))
