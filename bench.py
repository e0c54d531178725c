#!/usr/bin/env python3
"""Benchmark of the qasMHD hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload turb2d]

Workload (default): 2D periodic decaying MHD, 1024^2 interior cells on
[0, 2pi]^2, gam = 5/3, rho = gam^2, p = gam, velocity and Az from 16 seeded
Fourier modes (|k| <= 4) with rms|v| = rms|B| = 1; HLLD, WCNS-4, CUCT CT,
SSP-RK3, CFL 0.4 with the reference's dt schedule (dt from setdt every 2nd
step).  One "step" is one full RK3 step (3 stages with per-stage ghost fill)
plus the CFL reduction every 2nd step, on a state already resident in HBM;
the metric is interior zone-cycles per second.

Also reported, on one JSON line:
  mixed         FP64 vs mixed precision (FP32 reconstruction / flux, FP64
                state, RK and pressure recovery) on the low-Mach
                Kelvin-Helmholtz workload at 2048^2: both rates and the L1
                difference of density and pressure
  batch         64 periodic instances of uneven sizes (128-512 square) in one
                batch (one launch per stage kernel for all) against the same
                instances run one after another
  adjoint       time of the hand-written discrete adjoint over 20 steps of a
                256^2 periodic run against the plain forward of those steps
  dissipative   2D reconnection 4096 x 512 with localised viscosity and
                resistivity: ideal step plus sub-cycled visco-resistive step,
                one dissipative call alone, and the reference's own one
  gravity       2D Rayleigh-Taylor 512 x 4096 with gravity, and the
                reference's own RTI driver on this host
  inverse       64 batched instances of the velocity-mode inversion: the
                batched forward, the adjoint gradient of all 512 amplitudes
                and a finite-difference check
  turb3d        3D periodic MHD turbulence at 512^3 (HLLD, WCNS-4, 3D CT,
                RK3; initial state built on the device with torch): rate over
                20 steps and the roofline of its dominant sweep
  roofline      the dominant kernel's algorithmic work per launch (flops and
                compulsory bytes from the library's work model) over its
                average launch time, measured with HIP events on the
                library's stream during the timed region
  cpu_baseline  the reference's own 2D driver (oracle/_ref, built from the
                reference sources) timed on this host over a bounded sample
                of the same workload; the C oracle when that is not built
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-cycles/s at 1 GPU, FP64 vs mixed at stated L1 err; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from qasmhd_amd import problems
    if name == "turb2d":
        return problems.turbulence2d(n=1024, seed=2), "2d_decaying_mhd_1024x1024_hlld_wcns4_ct_rk3"
    if name == "turb2d_512":
        return problems.turbulence2d(n=512, seed=2), "2d_decaying_mhd_512x512_hlld_wcns4_ct_rk3"
    if name == "khi":
        return problems.kelvin_helmholtz(n=2048), "2d_lowmach_khi_2048x2048_mlau_wcns4_ct_rk3"
    if name.startswith("turb3d"):
        n = int(name[6:] or 512)
        return problems.turbulence3d(n=n, device="cuda:0"), "3d_mhd_turbulence_%dx%dx%d_hlld_wcns4_ct_rk3" % (n, n, n)
    if name == "brio_wu":
        return problems.brio_wu(1024), "1d_brio_wu_1024_hlld_wcns4_rk3"
    raise SystemExit("unknown workload %s" % name)


def recorded_traffic(workload_name, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (profiles/*_traffic.md, written from two rocprofv3 --pmc
    passes, FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration)
    and where they come from (file, and the code commit the counters were
    taken at, as the file states it); (None, None) when there is none.  The
    counters cannot be read from inside this process, so the figure is the
    committed measurement, not one of this run."""
    import glob
    import re
    tag = {"2d_decaying_mhd_1024x1024_hlld_wcns4_ct_rk3": "stage2d_turb1024_fp64"}.get(workload_name)
    if tag is None:
        return None, None
    for path in glob.glob(os.path.join(ROOT, "profiles", tag + "_traffic.md")):
        with open(path) as f:
            text = f.read()
        m = re.search(r"%s=(\d+)" % kernel, text)
        c = re.search(r"code commit: ([0-9a-f]+)", text)
        if m:
            return float(m.group(1)), {"file": os.path.relpath(path, ROOT),
                                       "code_commit": c.group(1) if c else None,
                                       "how": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, per dispatch, scaled by "
                                              "a device copy of known size; read from the committed summary"}
    return None, None


def host_info():
    """The host CPU the baselines ran on (lscpu) and the OpenMP placement."""
    import subprocess
    info = {"visible_logical_cpus": len(os.sched_getaffinity(0)),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v.strip()
    except Exception as e:  # reported, not required
        info["lscpu_error"] = str(e)
    return info


def cpu_threads():
    """OpenMP threads of the CPU baselines: OMP_NUM_THREADS when set (the
    GPU host sets it to its CPU share), else the visible CPUs up to 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return min(16, len(os.sched_getaffinity(0)))


def cpu_baseline(p, nsteps):
    """The reference's own driver on this host (kind 'reference'), else the C
    oracle (kind 'port'); same initial state, first `nsteps` steps."""
    import numpy as np
    from oracle import build as obuild
    threads = cpu_threads()
    zones = int(np.prod(p.interior))
    if p.ndim == 2:
        nx, ny = p.interior[1], p.interior[0]
        binary = os.path.join(obuild.REFOUT, obuild.ref_binary_name(2, nx, ny, p.rmn, p.odr, p.r_k, p.ctw, p.cfl))
        if os.path.exists(binary):
            _, _, el = obuild.run_ref(binary, p.state, p.dx, p.gam, nsteps, 1, dnx=list(p.dn[0]),
                                      dny=list(p.dn[1]), dy=p.dy, threads=threads, timing=True)
            return {"value": zones * nsteps / el, "unit": "zone-cycles/s", "cores": threads, "kind": "reference",
                    "sample": "first %d steps of the same %dx%d run (reference MHD2D::ideal/bound/setdt with its "
                              "setdt schedule, g++ -O2 -fopenmp, OMP_NUM_THREADS=%d, OMP_PROC_BIND=%s, OMP_PLACES=%s), "
                              "%.2f s wall for the step loop" % (nsteps, nx, ny, threads, os.environ.get("OMP_PROC_BIND"),
                                                                 os.environ.get("OMP_PLACES"), el)}
    from oracle import pyoracle
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    cfg = pyoracle.make_cfg(p.state.shape[1:], p.dx, p.dy, p.dz, gam=p.gam, cfl=p.cfl, rmn=p.rmn, odr=p.odr,
                            r_k=p.r_k, ctw=p.ctw, dn=p.dn)
    st = p.state.copy()
    t0 = time.perf_counter()
    pyoracle.run(cfg, st, nsteps, 1)
    el = time.perf_counter() - t0
    return {"value": zones * nsteps / el, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "sample": "first %d steps of the same run (C oracle, OpenMP %d threads), %.2f s" % (nsteps, threads, el)}


def brio_wu_line(cpu=True):
    """The 1D Brio-Wu shock tube (1024 cells, gam 2, HLLD, WCNS-4, RK3, CFL
    0.4, open ends) to t = 0.1 at the fixed dt of the initial setdt, as the
    reference's 1D driver runs it (exec_(0), 1D/SERIAL/shock/main.cpp:9):
    GPU zone-cycles/s over the whole run, and the reference's own serial 1D
    driver (oracle/_ref) over the same steps on one host thread."""
    from qasmhd_amd import capi, problems
    p = problems.brio_wu(1024)
    zones = int(np.prod(p.interior))
    with capi.solver_for(p) as s:
        s.upload(p.state)
        dt0 = s.cfl_dt()
        nmax = int(0.1 / dt0 + 0.5)  # mhd1d_class.cpp: nmax = (int)(tmax/dt + 0.5)
        s.run(4, flg=0, dt=dt0, history=False)  # warm-up
        s.upload(p.state)
        s.reset_counter()
        s.sync()
        t0 = time.perf_counter()
        s.run(nmax, flg=0, dt=dt0, history=False)
        s.sync()
        el = time.perf_counter() - t0
        out = s.download()
    line = {"workload": "1d_brio_wu_1024_hlld_wcns4_rk3", "steps": nmax, "dt": dt0, "value": zones * nmax / el,
            "unit": "zone-cycles/s", "ms_per_step": el / nmax * 1e3, "finite": bool(np.isfinite(out).all())}
    if cpu:
        try:
            from oracle import build as obuild
            binary = os.path.join(obuild.REFOUT, obuild.ref_binary_name(1, 1024, 1, p.rmn, p.odr, p.r_k, 1, p.cfl))
            if os.path.exists(binary):
                ref, _, el_ref = obuild.run_ref(binary, p.state, p.dx, p.gam, nmax, 0, dt=dt0, dnx=list(p.dn[0]),
                                                threads=1, timing=True)
                line["cpu_baseline"] = {"value": zones * nmax / el_ref, "unit": "zone-cycles/s", "cores": 1,
                                        "kind": "reference",
                                        "sample": "the whole run, %d steps (reference MHD1D::ideal/bound, g++ -O2 "
                                                  "-fopenmp, 1 thread: the serial 1D path), %.3f s" % (nmax, el_ref)}
                line["equals_reference"] = bool(np.array_equal(ref, out))
        except Exception as e:  # reported, not required
            _log("bench: reference 1D timing failed: %s" % e)
    return line


def batch_cpu_reference(sizes, steps=4):
    """The reference's 2D driver on this host for the batched workload: its
    grid size is a compile-time macro and its step keeps static state
    (mhd2d_solve.cpp:8), so each instance is one process; every distinct
    size is timed over `steps` steps of its first instance and the batch
    rate is total zones over the summed per-instance step times."""
    from qasmhd_amd import problems
    from oracle import build as obuild
    threads = cpu_threads()
    per = {}
    for n in sorted(set(int(v) for v in sizes)):
        binary = os.path.join(obuild.REFOUT, obuild.ref_binary_name(2, n, n, 1, 4, 3, 1, 0.4))
        if not os.path.exists(binary):
            return None
        p = problems.turbulence2d(n=n, seed=5, vrms=0.3, brms=0.3)
        _, _, el = obuild.run_ref(binary, p.state, p.dx, p.gam, steps, 1, dnx=list(p.dn[0]), dny=list(p.dn[1]),
                                  dy=p.dy, threads=threads, timing=True)
        per[n] = el / steps
    zones = sum(int(n) * int(n) for n in sizes)
    t_step = sum(per[int(n)] for n in sizes)
    return {"value": zones / t_step, "unit": "zone-cycles/s", "cores": threads, "kind": "reference",
            "sample": "each distinct instance size (%s) for %d steps of the reference MHD2D (g++ -O2 -fopenmp, "
                      "OMP_NUM_THREADS=%d), one process per instance, instances one after another"
                      % (",".join(str(n) for n in sorted(per)), steps, threads),
            "seconds_per_step": {str(k): v for k, v in per.items()}}


def cpu_port_3d(n, steps, seed=4):
    """The oracle (this build's C restatement of the reference's 3D step,
    OpenMP) on the 3D turbulence set-up at n^3 for `steps` steps of the
    setdt schedule: the 3D reference driver needs MPI, which this host
    lacks.  The initial state is built on the device with torch and
    copied to the host."""
    import torch
    from oracle import pyoracle
    from qasmhd_amd import problems
    threads = cpu_threads()
    p = problems.turbulence3d(n=n, seed=seed, device="cuda:0")
    st = p.state.cpu().numpy().copy()
    del p.state
    torch.cuda.empty_cache()
    cfg = pyoracle.make_cfg(st.shape[1:], p.dx, p.dy, p.dz, gam=p.gam, cfl=p.cfl, rmn=p.rmn, odr=p.odr, r_k=p.r_k,
                            ctw=p.ctw)
    t0 = time.perf_counter()
    pyoracle.run(cfg, st, steps, 1)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * steps / el, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "sample": "%d steps of the same set-up at %d^3 (C oracle: the reference's 3D step restated, gcc -O2 "
                      "-fopenmp, OMP_NUM_THREADS=%d, OMP_PROC_BIND=%s, OMP_PLACES=%s; the reference's 3D driver "
                      "needs MPI), %.2f s" % (steps, n, threads, os.environ.get("OMP_PROC_BIND"),
                                             os.environ.get("OMP_PLACES"), el),
            "finite": bool(np.isfinite(st).all())}


def _interior(a, ndim):
    g = 4
    return a[(Ellipsis,) + (slice(g, -g),) * ndim]


def mixed_line(n, steps):
    """FP64 vs mixed precision on the low-Mach Kelvin-Helmholtz workload
    (MLAU, WCNS-4, CUCT, RK3): zone-cycles/s of both at the same fixed dt
    and the L1 difference of density and pressure after `steps` steps."""
    import numpy as np
    from qasmhd_amd import capi, problems
    p = problems.kelvin_helmholtz(n=n)
    zones = int(np.prod(p.interior))
    with capi.solver_for(p) as s:
        s.upload(p.state)
        dt = s.cfl_dt()
    out, rate = {}, {0: 0.0, 1: 0.0}
    solvers = {prec: capi.solver_for(p, precision=prec) for prec in (0, 1)}
    for s in solvers.values():
        s.upload(p.state)
        s.run(4, flg=0, dt=dt, history=False)  # warm-up
    for trial in range(2):  # alternate the two precisions; best of two each
        for prec in (0, 1):
            s = solvers[prec]
            s.upload(p.state)
            s.reset_counter()
            s.sync()
            t0 = time.perf_counter()
            s.run(steps, flg=0, dt=dt, history=False)
            s.sync()
            el = time.perf_counter() - t0
            out[prec] = s.download()
            rate[prec] = max(rate[prec], zones * steps / el)
    for s in solvers.values():
        s.close()
    gam = p.gam

    def pres(st):
        ro = st[0]
        v2 = (st[1] ** 2 + st[2] ** 2 + st[3] ** 2) / ro
        bx = 0.5 * (st[4] + np.roll(st[4], -1, axis=-1))
        by = 0.5 * (st[5] + np.roll(st[5], -1, axis=-2))
        return (gam - 1) * (st[7] - 0.5 * (v2 + bx ** 2 + by ** 2 + st[6] ** 2))
    a, b = out[0], out[1]
    l1r = float(np.mean(np.abs(_interior(b[0] - a[0], 2))))
    l1p = float(np.mean(np.abs(_interior(pres(b) - pres(a), 2))))
    return {"workload": "2d_lowmach_khi_%dx%d_mlau_wcns4_ct_rk3" % (n, n), "steps": steps, "dt": dt,
            "fp64_value": rate[0], "mixed_value": rate[1], "speedup": rate[1] / rate[0], "l1_rho": l1r,
            "l1_p": l1p, "finite": bool(np.isfinite(b).all())}


def fma64_line(p, wname, warmup, steps):
    """FP64 with fused multiply-adds in the sweeps (precision 2) against the
    bit-exact FP64 path on the same workload, both timed as the headline is
    (setdt schedule, `steps` steps after `warmup`), best of two
    alternating trials; relative L1 distance of every variable over the
    interior after warmup + steps steps and the final dt."""
    import numpy as np
    from qasmhd_amd import capi
    zones = int(np.prod(p.interior))
    rate, out, dts = {0: 0.0, 2: 0.0}, {}, {}
    solvers = {prec: capi.solver_for(p, precision=prec) for prec in (0, 2)}
    for trial in range(2):
        for prec in (0, 2):
            s = solvers[prec]
            upload_problem(s, p)
            s.reset_counter()
            s.run(warmup, flg=1, history=False)
            s.sync()
            t0 = time.perf_counter()
            _, dts[prec] = s.run(steps, flg=1, history=False)
            s.sync()
            rate[prec] = max(rate[prec], zones * steps / (time.perf_counter() - t0))
            out[prec] = s.download()
    for s in solvers.values():
        s.close()
    a, b = out[0], out[2]
    # a vector component against the mean L1 norm of its vector's components
    groups = ((0,), (1, 2, 3), (1, 2, 3), (1, 2, 3), (4, 5, 6), (4, 5, 6), (4, 5, 6), (7,))
    norm = [float(np.abs(_interior(a[m], p.ndim)).sum()) for m in range(8)]
    e = [float(np.abs(_interior(b[m] - a[m], p.ndim)).sum()) / (sum(norm[k] for k in groups[m]) / len(groups[m]))
         for m in range(8)]
    return {"workload": wname, "steps": steps, "warmup": warmup, "fp64_exact_value": rate[0],
            "fp64_fma_value": rate[2], "speedup": rate[2] / rate[0],
            "rel_l1_max_over_variables": max(e), "rel_l1": e,
            "dt_rel_diff": abs(dts[2] / dts[0] - 1.0), "finite": bool(np.isfinite(b).all()),
            "note": "precision 2: the same FP64 sweeps built with FMA contraction; distance to the bit-exact "
                    "(reference-identical) path after warmup + steps steps, bar 1e-12"}


def batch_line(ninst, steps, seed=5):
    """Batched domains (qas_batch_run): `ninst` periodic 2D instances of
    square sizes drawn (numpy PCG64(seed)) from {128, 192, 256, 384, 512},
    rms Mach 0.3, HLLD / WCNS-4 / CUCT / RK3, `steps` steps in one batch
    (one launch per stage kernel for all instances, two steps per graph
    replay; the launch shape picked by the batch's autotuner on the first
    run); against the same instances run one after another with qas_run."""
    import numpy as np
    from qasmhd_amd import capi, problems
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.choice([128, 192, 256, 384, 512], size=ninst)
    ps = [problems.turbulence2d(n=int(n), seed=seed * 1000 + i, vrms=0.3, brms=0.3) for i, n in enumerate(sizes)]
    zones = int(sum(int(np.prod(p.interior)) for p in ps))
    solvers = [capi.solver_for(p) for p in ps]
    for s, p in zip(solvers, ps):
        s.upload(p.state)
    with capi.Batch(solvers) as b:
        # launch-shape tuning, capture of the whole run as one graph, warm-up
        b.run(steps, flg=1, history=False)
        shape = b.shape()
        for s, p in zip(solvers, ps):
            s.upload(p.state)
            s.reset_counter()
        t0 = time.perf_counter()
        b.run(steps, flg=1, history=False)
        for s in solvers:
            s.sync()
        t_batch = time.perf_counter() - t0
    for s, p in zip(solvers, ps):
        s.upload(p.state)
        s.reset_counter()
        s.run(2, flg=1, history=False)
        s.upload(p.state)
        s.reset_counter()
        s.sync()
    t0 = time.perf_counter()
    for s in solvers:
        s.run(steps, flg=1, history=False)
    for s in solvers:
        s.sync()
    t_seq = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return {"workload": "%d_instances_2d_periodic_sizes_128-512_mach0.3_hlld_wcns4_ct_rk3" % ninst, "steps": steps,
            "zones": zones, "batched_value": zones * steps / t_batch, "sequential_value": zones * steps / t_seq,
            "speedup": t_seq / t_batch,
            "shape": {"order": ["as_given", "largest_first"][shape[0]], "x_waves_per_workgroup": shape[1],
                      "elementwise_workgroup": shape[2]},
            "sizes": [int(v) for v in sizes]}


def adjoint_line(n, steps, rmn=1):
    """Discrete adjoint (qas_adj_forward / qas_adj_backward) on the 2D
    periodic workload at rms Mach 0.3 (HLLD, or MLAU with rmn 3; WCNS-4,
    CUCT, RK3): time of the backward pass over `steps` steps against the
    plain forward of the same steps at the same dt values."""
    import numpy as np
    from qasmhd_amd import capi, problems
    p = problems.turbulence2d(n=n, seed=5, vrms=0.3, brms=0.3)
    rng = np.random.default_rng(5)
    W = rng.standard_normal(p.state.shape)
    with capi.solver_for(p, rmn=rmn) as s:
        s.upload(p.state)
        dts, _ = s.run(steps, flg=1)
        # plain forward at the recorded step sizes
        s.upload(p.state)
        s.bound()
        s.sync()
        t0 = time.perf_counter()
        for dt in dts:
            s.step(dt)
        s.sync()
        t_fwd = time.perf_counter() - t0
        s.upload(p.state)
        s.sync()
        t0 = time.perf_counter()
        s.adj_forward(dts)
        t_rec = time.perf_counter() - t0
        s.adj_backward(W)  # first call allocates and warms up
        t0 = time.perf_counter()
        g = s.adj_backward(W)
        t_bwd_host = time.perf_counter() - t0
        with capi.DeviceArray(W.shape) as dg, capi.DeviceArray(W.shape) as di:
            dg.upload(W)
            t0 = time.perf_counter()
            s.adj_backward_device(dg, di)
            t_bwd = time.perf_counter() - t0
            gd = di.download()
            # the same with the forward keeping every stage's intermediates
            # (qas_adj_keep_stages): the backward skips the recompute; the
            # first recorded forward allocates the store, the second is timed
            s.adj_keep_stages(True)
            t_keep = []
            for _ in range(2):
                s.upload(p.state)
                s.sync()
                t0 = time.perf_counter()
                s.adj_forward(dts)
                t_keep.append(time.perf_counter() - t0)
            s.adj_backward_device(dg, di)
            s.sync()
            t0 = time.perf_counter()
            s.adj_backward_device(dg, di)
            t_bwd_k = time.perf_counter() - t0
            gk = di.download()
            s.adj_keep_stages(False)
    nd = int(np.prod(p.state.shape[1:]))
    return {"workload": "2d_decaying_mhd_%dx%d_mach0.3_%s_wcns4_ct_rk3" % (n, n, "mlau" if rmn == 3 else "hlld"),
            "steps": steps,
            "forward_s": t_fwd, "forward_with_checkpoints_s": t_rec, "backward_s": t_bwd,
            "backward_with_host_copies_s": t_bwd_host, "ratio": t_bwd / t_fwd,
            "gradient_ratio": (t_rec + t_bwd) / t_fwd,
            "finite": bool(np.isfinite(g).all()), "device_equals_host": bool(np.array_equal(g, gd)),
            "kept_stages": {"forward_s": t_keep[1], "backward_s": t_bwd_k, "ratio": t_bwd_k / t_fwd,
                            "gradient_ratio": (t_keep[1] + t_bwd_k) / t_fwd,
                            "bytes_per_step": 8 * 54 * nd, "checkpoint_bytes_per_step": 8 * 8 * nd,
                            "equals_recompute": bool(np.array_equal(gk, gd))},
            "note": "ratio = backward / forward, gradient buffers resident in HBM (qas_adj_backward_device); "
                    "backward = per step: 3 recomputed forward stages + 3 stage adjoints; gradient_ratio = "
                    "(checkpointed forward + backward) / forward; kept_stages: the forward keeps every "
                    "stage's intermediates (qas_adj_keep_stages) and the backward skips the recompute"}


def kernel_roofline(s, lib, sweeps, wname):
    """Roofline object of the dominant sweep kernel: its algorithmic flops and
    compulsory bytes per launch (qas_kernel_model) over its average launch
    time from the HIP events recorded on the library's stream."""
    import ctypes as C
    best = None
    for k in sweeps:
        tot, n = C.c_double(), C.c_longlong()
        s._check(lib.qas_prof_query(s.ctx, k, C.byref(tot), C.byref(n)), "qas_prof_query")
        fl, by = C.c_double(), C.c_double()
        s._check(lib.qas_kernel_model(s.ctx, k, C.byref(fl), C.byref(by)), "qas_kernel_model")
        if n.value == 0:
            continue
        avg_s = tot.value / n.value * 1e-3
        rec = dict(kid=k, total_ms=tot.value, launches=n.value, avg_s=avg_s, flops=fl.value, bytes=by.value)
        if best is None or tot.value > best["total_ms"]:
            best = rec
    names = {1: "sweep_x", 2: "sweep_y", 3: "sweep_z"}
    if best is None:
        return None
    t_flop = best["flops"] / (FP64_PEAK_TFLOPS * 1e12)
    t_hbm = best["bytes"] / (HBM_PEAK_GBS * 1e9)
    tflops = best["flops"] / best["avg_s"] / 1e12
    gbs = best["bytes"] / best["avg_s"] / 1e9
    if t_hbm > t_flop:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    else:
        roof = {"bound": "mfma", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tflops / FP64_PEAK_TFLOPS}
    traffic, tsrc = recorded_traffic(wname, names[best["kid"]])
    roof.update({"traffic": traffic, "traffic_source": tsrc, "kernel": names[best["kid"]],
                 "avg_launch_us": best["avg_s"] * 1e6, "launches": best["launches"],
                 "alg_flops_per_launch": best["flops"], "alg_bytes_per_launch": best["bytes"],
                 "fp64_tflops": tflops, "hbm_gbs": gbs, "fp64_frac": tflops / FP64_PEAK_TFLOPS,
                 "hbm_frac": gbs / HBM_PEAK_GBS,
                 "compute": "FP64 VALU (no MFMA on this stencil path); peak = MI355X FP64 78.6 TFLOP/s"})
    return roof


def upload_problem(s, p):
    """Copy a problem's state in: a numpy array from the host, a torch tensor
    (device-built initial condition) device to device."""
    if isinstance(p.state, np.ndarray):
        s.upload(p.state)
        return
    import torch
    torch.cuda.synchronize()
    s.upload_device(p.state.data_ptr())
    s.sync()


def state_finite(s, p):
    """All state values finite (checked on the device for a device problem)."""
    if isinstance(p.state, np.ndarray):
        return bool(np.isfinite(s.download()).all())
    import torch
    out = torch.empty(tuple(p.state.shape), dtype=torch.float64, device="cuda:0")
    s.download_device(out.data_ptr())
    s.sync()
    return bool(torch.isfinite(out).all())


def enable_sweep_prof(s, lib, ndim, nstages):
    sweeps = [1, 2, 3][:ndim]
    mask = 0
    for k in sweeps:
        mask |= 1 << k
    s._check(lib.qas_prof_enable(s.ctx, mask, nstages * len(sweeps) + 16), "qas_prof_enable")
    return sweeps


def inverse_line(ninst, steps, seed=5, nfd=2):
    """Velocity-mode inversion on `ninst` batched 2D instances (sizes drawn
    from {128, 192, 256, 384, 512}, M = 0.3, 8 vy-mode amplitudes each,
    qasmhd_amd/inverse.py): the batched forward that fixes the dt schedule,
    the loss gradient of all instances by the discrete adjoint, and a
    central-difference check (h = 1e-6) of `nfd` amplitudes of the first
    instance."""
    from qasmhd_amd import inverse
    sizes = inverse.config_sizes(ninst, seed=seed)
    inv = inverse.VyModeInversion(sizes, steps=steps, seed=seed)
    zones = int(sum(n * n for n in sizes))
    inv.schedule(warm=True)  # fixes the schedule
    t_fwd = inv.batch_seconds
    inv.targets()
    # first call: contexts, checkpoint room and launch shape; timed: the
    # second, as every further gradient of an iterative inversion runs
    grp = {}
    for groups in (1, 4):
        t0 = time.perf_counter()
        inv.loss_and_grad_batched(inv.a0, groups=groups)
        t_first = time.perf_counter() - t0
        t0 = time.perf_counter()
        L, G = inv.loss_and_grad_batched(inv.a0, groups=groups)
        grp[groups] = (time.perf_counter() - t0, t_first)
        inv.close()
    t_bat, t_grad_first = min(grp.values())
    # batched recorded forward, then each instance's own backward (4 groups
    # on 4 host threads), second call timed
    inv.loss_and_grad_batched(inv.a0, groups=4, backward="per_instance")
    t0 = time.perf_counter()
    Lh, Gh = inv.loss_and_grad_batched(inv.a0, groups=4, backward="per_instance")
    t_hyb = time.perf_counter() - t0
    inv.close()
    t0 = time.perf_counter()
    L4, G4 = inv.loss_and_grad(inv.a0)
    t_grad_thr_fresh = time.perf_counter() - t0
    # per instance with kept contexts (second call timed, as for the batches)
    inv.loss_and_grad(inv.a0, keep=True)
    t0 = time.perf_counter()
    Lk, Gk = inv.loss_and_grad(inv.a0, keep=True)
    t_grad_thr = time.perf_counter() - t0
    inv.close()
    paths = {"batched": t_bat, "batched_forward_per_instance_backward": t_hyb, "per_instance_threads": t_grad_thr}
    best = min(paths, key=paths.get)
    errs = []
    scale = float(np.abs(G[0]).max())
    for m in range(nfd):
        fd = inv.fd_grad(0, inv.a0[0], m, h=1e-6)
        errs.append(abs(fd - G[0, m]) / scale)
    return {"workload": "%d_instances_2d_vy_mode_inversion_sizes_128-512_mach0.3_hlld_wcns4_ct_rk3" % ninst,
            "steps": steps, "zones": zones, "params": int(G.size), "loss": L,
            "batched_forward_s": t_fwd, "batched_forward_value": zones * steps / t_fwd,
            "gradient_s": paths[best], "ratio": paths[best] / t_fwd, "gradient_path": best,
            "fd_max_rel_err": max(errs) if errs else None, "fd_checked": nfd, "finite": bool(np.isfinite(G).all()),
            "gradient_batched_s": t_bat, "gradient_batched_first_call_s": t_grad_first,
            "gradient_batched_s_by_groups": {str(k): v[0] for k, v in grp.items()},
            "gradient_per_instance_threads_s": t_grad_thr,
            "gradient_per_instance_threads_fresh_contexts_s": t_grad_thr_fresh,
            "gradient_batched_forward_per_instance_backward_s": t_hyb,
            "batched_equals_per_instance": bool(L == L4 and np.array_equal(G, G4) and Lh == L4
                                                and np.array_equal(Gh, G4) and Lk == L4 and np.array_equal(Gk, G4)),
            "note": "gradient_s = the fastest of three paths with bit-identical results: batched adjoint "
                    "(second call, contexts and checkpoint room kept from the first; instances in 1 or 4 "
                    "size-balanced batches on as many host threads: qas_batch_adj_forward records the forward "
                    "of a batch in one batched run, qas_batch_adj_backward runs per step the 3 recomputed "
                    "stages and 3 stage adjoints with one launch per kernel for the batch), the batched "
                    "recorded forward followed by each instance's own backward (4 batches on 4 host threads, "
                    "second call), and per instance (own context and stream, 4 host threads; contexts kept from a "
                    "first call, the fresh-context time reported beside it); all include host "
                    "copies of states and gradients; forward = one batched run of all instances (graph replay)"}


def gravity_line(nx, ny, steps, cpu=True):
    """2D magnetised Rayleigh-Taylor with gravity along y (the reference's
    RTI set-up: LHLLD, WCNS-4, RK3, CUCT, heavy over light fluid) at
    nx x ny: zone-cycles/s of exec_ steps with the gravity terms, and the
    reference's own RTI driver on this host over a bounded sample."""
    from qasmhd_amd import capi, problems
    p, phi = problems.rayleigh_taylor2d(nx=nx, ny=ny)
    zones = nx * ny
    with capi.solver_for(p) as s:
        s.upload(p.state)
        s.set_gravity(phi)
        s.run(2, flg=1, history=False)  # warm-up
        s.upload(p.state)
        s.reset_counter()
        s.sync()
        t0 = time.perf_counter()
        _, dt_now = s.run(steps, flg=1, history=False)
        s.sync()
        el = time.perf_counter() - t0
        finite = bool(np.isfinite(s.download()).all())
    out = {"workload": "2d_rayleigh_taylor_%dx%d_lhlld_wcns4_ct_rk3_gravity" % (nx, ny), "steps": steps,
           "value": zones * steps / el, "unit": "zone-cycles/s", "ms_per_step": el / steps * 1e3,
           "dt_final": dt_now, "finite": finite}
    if cpu:
        try:
            from oracle import build as obuild
            binary = os.path.join(obuild.REFOUT, obuild.ref_binary_name(2, nx, ny, p.rmn, p.odr, p.r_k, p.ctw, p.cfl,
                                                                        gravity=True))
            if os.path.exists(binary):
                threads = cpu_threads()
                k = 4
                _, _, el_ref = obuild.run_ref(binary, p.state, p.dx, p.gam, k, 1, dnx=list(p.dn[0]),
                                              dny=list(p.dn[1]), dy=p.dy, threads=threads, timing=True, phi=phi)
                out["cpu_reference"] = {"value": zones * k / el_ref, "unit": "zone-cycles/s", "cores": threads,
                                        "kind": "reference",
                                        "sample": "first %d steps of the same run (reference RTI MHD2D, g++ -O2 "
                                                  "-fopenmp), %.2f s" % (k, el_ref)}
        except Exception as e:  # reported, not required
            _log("bench: reference RTI timing failed: %s" % e)
    return out


def dissipative_line(nx, ny, steps, cpu=True):
    """2D Harris-sheet reconnection with localised viscosity and resistivity
    (the reference's dissipative MRX set-up: LHLLD, ODR 2, RK3, CUCT,
    Reynolds numbers 200) at nx x ny: zone-cycles/s of DMHD2D::exec_ steps
    (ideal(dt), then ncmax sub-steps of the visco-resistive dsptv), the time
    of one dsptv call alone, and the reference's own dsptv on this host over
    a bounded sample."""
    from qasmhd_amd import capi, problems
    p, nu, eta, dcmax = problems.reconnection2d(nx=nx, ny=ny)
    zones = nx * ny
    dr = min(p.dx, p.dy)
    with capi.solver_for(p) as s:
        s.upload(p.state)
        s.set_dissipation(nu, eta, dcmax)
        s.run(2, flg=1, history=False)  # warm-up
        s.upload(p.state)
        s.reset_counter()
        s.sync()
        t0 = time.perf_counter()
        hist, dt_now = s.run(steps, flg=1)
        s.sync()
        el = time.perf_counter() - t0
        ncmax = [1 + int(2 * (dcmax * d) / (dr * dr)) for d in hist]
        dt_sub = float(hist[-1]) / ncmax[-1]
        nd = 20
        s.sync()
        t0 = time.perf_counter()
        for _ in range(nd):
            s.dsptv(dt_sub)
        s.sync()
        t_dsp = (time.perf_counter() - t0) / nd
        finite = bool(np.isfinite(s.download()).all())
    out = {"workload": "2d_reconnection_%dx%d_lhlld_odr2_ct_rk3_visco_resistive" % (nx, ny), "steps": steps,
           "value": zones * steps / el, "unit": "zone-cycles/s", "ms_per_step": el / steps * 1e3,
           "ncmax": [min(ncmax), max(ncmax)], "dsptv_ms": t_dsp * 1e3, "dsptv_zones_per_s": zones / t_dsp,
           "dcmax": dcmax, "finite": finite}
    if cpu:
        try:
            from oracle import build as obuild
            binary = os.path.join(obuild.REFOUT, obuild.ref_dsptv_name(nx, ny, p.rmn, p.odr, p.r_k, p.ctw))
            if os.path.exists(binary):
                threads = cpu_threads()
                k = 4
                _, el_ref = obuild.run_ref_dsptv(binary, p.state, nu, eta, p.dx, p.dy, p.gam, dt_sub, nsub=k,
                                                 dnx=list(p.dn[0]), dny=list(p.dn[1]), threads=threads,
                                                 timing=True)
                out["cpu_reference"] = {"dsptv_zones_per_s": zones * k / el_ref, "cores": threads,
                                        "kind": "reference",
                                        "sample": "%d calls of bound + DMHD2D::dsptv (g++ -O2 -fopenmp), %.2f s"
                                                  % (k, el_ref)}
        except Exception as e:  # reported, not required
            _log("bench: reference dsptv timing failed: %s" % e)
    return out


def mixed3d_line(p, steps):
    """FP64 vs mixed precision on a device-built 3D problem: both rates over
    `steps` steps at the same fixed dt, and the L1 difference of density and
    pressure of the final states (computed on the device)."""
    import torch
    from qasmhd_amd import capi
    zones = int(np.prod(p.interior))
    shape = tuple(p.state.shape)
    rate, fin, dt = {}, {}, None
    for prec in (0, 1):
        with capi.solver_for(p, precision=prec) as s:
            upload_problem(s, p)
            if dt is None:
                dt = s.cfl_dt()
            s.run(2, flg=0, dt=dt, history=False)  # warm-up
            upload_problem(s, p)
            s.reset_counter()
            s.sync()
            t0 = time.perf_counter()
            s.run(steps, flg=0, dt=dt, history=False)
            s.sync()
            rate[prec] = zones * steps / (time.perf_counter() - t0)
            f = torch.empty(shape, dtype=torch.float64, device="cuda:0")
            s.download_device(f.data_ptr())
            s.sync()
            fin[prec] = f
    gam = p.gam
    I = (slice(4, -4),) * 3

    def pres(st):
        ro = st[0]
        v2 = (st[1] ** 2 + st[2] ** 2 + st[3] ** 2) / ro
        bx = 0.5 * (st[4] + torch.roll(st[4], -1, dims=2))
        by = 0.5 * (st[5] + torch.roll(st[5], -1, dims=1))
        bz = 0.5 * (st[6] + torch.roll(st[6], -1, dims=0))
        return (gam - 1) * (st[7] - 0.5 * (v2 + bx ** 2 + by ** 2 + bz ** 2))
    a, b = fin[0], fin[1]
    l1r = float(torch.mean(torch.abs(b[0][I] - a[0][I])))
    l1p = float(torch.mean(torch.abs(pres(b)[I] - pres(a)[I])))
    finite = bool(torch.isfinite(b).all())
    del fin, a, b
    torch.cuda.empty_cache()
    return {"steps": steps, "dt": dt, "fp64_value": rate[0], "mixed_value": rate[1], "speedup": rate[1] / rate[0],
            "l1_rho": l1r, "l1_p": l1p, "finite": finite}


def turb3d_child(n, steps, warmup=2, cpu_n=0):
    """3D periodic MHD turbulence at n^3 (HLLD, WCNS-4, 3D CT, RK3, CFL 0.4;
    initial state built on the device with torch): this script's own main
    line for that workload, run as a child process so that this process
    keeps its HIP runtime free of torch's; with cpu_n, also the C oracle's
    rate on the same set-up at cpu_n^3."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "turb3d%d" % n, "--steps", str(steps),
           "--warmup", str(warmup), "--no-cpu", "--no-extra", "--mixed3d-steps", str(steps),
           "--cpu3d-n", str(cpu_n)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("3D child exited with %d" % r.returncode)
    line = json.loads(r.stdout.strip().splitlines()[-1])
    keep = ("value", "unit", "steps", "warmup", "ms_per_step", "roofline", "mixed", "cpu_port")
    out = {"workload": line["config"]["workload"]}
    out.update({k: line[k] for k in keep if k in line})
    out.update({k: line["config"][k] for k in ("dt_final", "finite")})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="turb2d")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-precision and adjoint lines")
    ap.add_argument("--no-graphs", action="store_true",
                    help="every solver and batch launches every kernel from the host (qas_run's default)")
    ap.add_argument("--graphs", action="store_true", help="every solver and batch replays captured step-pair graphs")
    ap.add_argument("--precision", type=int, default=0,
                    help="0 FP64 (default), 1 mixed, 2 FP64 with FMA contraction, for the main line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy the 8 state arrays to a device buffer (8 copy "
                         "dispatches of nd doubles: the calibration of the memory counters)")
    ap.add_argument("--mixed-n", type=int, default=2048)
    ap.add_argument("--mixed-steps", type=int, default=200)
    ap.add_argument("--adj-n", type=int, default=256)
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--adj-steps", type=int, default=20)
    ap.add_argument("--adj-rmn", type=int, default=1, help="Riemann solver of the adjoint line (1 HLLD, 3 MLAU)")
    ap.add_argument("--inverse-n", type=int, default=64, help="instances of the inversion line (0: skip it)")
    ap.add_argument("--dsp-nx", type=int, default=4096, help="dissipative line size along x (0: skip it)")
    ap.add_argument("--dsp-ny", type=int, default=512)
    ap.add_argument("--dsp-steps", type=int, default=100)
    ap.add_argument("--grav-nx", type=int, default=512, help="gravity line size along x (0: skip it)")
    ap.add_argument("--grav-ny", type=int, default=4096)
    ap.add_argument("--grav-steps", type=int, default=100)
    ap.add_argument("--inverse-steps", type=int, default=200)
    ap.add_argument("--turb3d-n", type=int, default=512, help="3D turbulence line size (0: skip it)")
    ap.add_argument("--turb3d-steps", type=int, default=20)
    ap.add_argument("--cpu3d-n", type=int, default=0,
                    help="with a 3D workload: the C oracle as CPU baseline on the same set-up at this size")
    ap.add_argument("--cpu3d-steps", type=int, default=1)
    ap.add_argument("--khi-cpu-steps", type=int, default=4)
    ap.add_argument("--no-brio-wu", dest="brio_wu", action="store_false", help="skip the 1D line")
    ap.add_argument("--no-fma64", action="store_true", help="skip the FMA-contracted FP64 line")
    ap.add_argument("--mixed3d-steps", type=int, default=0,
                    help="with a 3D workload: also FP64 vs mixed over this many fixed-dt steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the domain is one device-resident block); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    import ctypes as C
    if args.workload.startswith("turb3d"):
        import torch  # before the library: both then bind torch's HIP runtime
        torch.cuda.init()
    from qasmhd_amd import capi
    if args.no_graphs or args.graphs:
        capi.GRAPHS = bool(args.graphs) and not args.no_graphs

    p, wname = workload(args.workload)
    zones = int(np.prod(p.interior))
    _log("bench: workload %s, %d zones, %d warmup + %d timed steps" % (wname, zones, args.warmup, args.steps))
    lib = capi.load()
    lib.qas_prof_enable.argtypes = [C.c_void_p, C.c_uint, C.c_int]
    lib.qas_prof_query.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    lib.qas_kernel_model.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    s = capi.solver_for(p, precision=args.precision)
    upload_problem(s, p)
    s.run(args.warmup, flg=1, history=False)
    s.sync()
    # the timed region: K steps as a user runs them (qas_run's default launch
    # mode, no timing hooks), the state resident in HBM
    t0 = time.perf_counter()
    hist, dt_now = s.run(args.steps, flg=1, history=False)
    s.sync()
    el = time.perf_counter() - t0
    finite = state_finite(s, p)
    shape = dict(zip(("x_waves_per_workgroup", "elementwise_workgroup"), s.shape_now()))
    value = zones * args.steps / el
    _log("bench: %.3f s for %d steps -> %.4g zone-cycles/s (finite=%s)" % (el, args.steps, value, finite))
    # a second pass of K steps, continuing the run, with the sweep kernels
    # launched through hipExtLaunchKernel with start/stop events (each event
    # stamped by the kernel's own dispatch): per-launch kernel times for the
    # roofline.  Graph replay is off in this pass.
    sweeps = enable_sweep_prof(s, lib, p.ndim, p.r_k * args.steps)
    s.sync()
    t1 = time.perf_counter()
    s.run(args.steps, flg=1, history=False)
    s.sync()
    el_prof = time.perf_counter() - t1
    roof = kernel_roofline(s, lib, sweeps, wname)
    if roof is not None:
        roof["timing"] = ("HIP events of hipExtLaunchKernel (the kernel's own start and end) on the library's stream "
                          "over a second pass of %d steps right after the timed region (%.4f ms per step in that pass, "
                          "host-launched)" % (args.steps, el_prof / args.steps * 1e3))
    s._check(lib.qas_prof_enable(s.ctx, 0, 0), "qas_prof_enable")
    if args.calibrate_copy:
        out = s.download()
        hip = C.CDLL("libamdhip64.so.7")
        buf = C.c_void_p()
        if hip.hipMalloc(C.byref(buf), C.c_size_t(out.nbytes)) != 0:
            raise SystemExit("bench: hipMalloc for the copy calibration failed")
        s.download_device(buf.value)
        s.sync()
        hip.hipFree(buf)
        _log("bench: copied 8 x %d bytes device to device (counter calibration)" % (out.nbytes // 8))
    s.close()

    cpu = None
    if not args.no_cpu and p.ndim == 2:
        try:
            _log("bench: cpu baseline, %d steps" % args.cpu_steps)
            cpu = cpu_baseline(p, args.cpu_steps)
        except Exception as e:  # the baseline is reported, not required
            _log("bench: cpu baseline failed: %s" % e)

    extra = {}
    if not args.no_fma64 and p.ndim == 2 and args.precision == 0:
        try:
            _log("bench: FP64 with FMA contraction vs bit-exact, %d steps" % args.steps)
            extra["fma64"] = fma64_line(p, wname, args.warmup, args.steps)
        except Exception as e:
            _log("bench: fma64 line failed: %s" % e)
    if p.ndim == 3 and args.mixed3d_steps > 0 and not isinstance(p.state, np.ndarray):
        try:
            _log("bench: 3D FP64 vs mixed, %d steps" % args.mixed3d_steps)
            extra["mixed"] = mixed3d_line(p, args.mixed3d_steps)
        except Exception as e:
            _log("bench: 3D mixed line failed: %s" % e)
    if p.ndim == 3 and args.cpu3d_n > 0 and not isinstance(p.state, np.ndarray):
        try:
            _log("bench: 3D CPU port baseline, %d^3, %d steps" % (args.cpu3d_n, args.cpu3d_steps))
            extra["cpu_port"] = cpu_port_3d(args.cpu3d_n, args.cpu3d_steps)
        except Exception as e:
            _log("bench: 3D CPU port baseline failed: %s" % e)
    if not args.no_extra and p.ndim == 2 and args.workload.startswith("turb2d") and args.turb3d_n > 0:
        try:
            _log("bench: 3D turbulence, %d^3, %d steps" % (args.turb3d_n, args.turb3d_steps))
            extra["turb3d"] = turb3d_child(args.turb3d_n, args.turb3d_steps,
                                           cpu_n=0 if args.no_cpu else args.cpu3d_n or 256)
        except Exception as e:
            _log("bench: 3D line failed: %s" % e)
    if not args.no_extra and p.ndim == 2 and args.workload.startswith("turb2d"):
        if args.mixed_n > 0:
            try:
                _log("bench: mixed precision, %d^2, %d steps" % (args.mixed_n, args.mixed_steps))
                extra["mixed"] = mixed_line(args.mixed_n, args.mixed_steps)
            except Exception as e:
                _log("bench: mixed line failed: %s" % e)
            if not args.no_cpu and "mixed" in extra:
                try:
                    from qasmhd_amd import problems
                    _log("bench: cpu baseline of the KH workload, %d steps" % args.khi_cpu_steps)
                    extra["mixed"]["cpu_baseline"] = cpu_baseline(problems.kelvin_helmholtz(n=args.mixed_n),
                                                                  args.khi_cpu_steps)
                except Exception as e:
                    _log("bench: KH cpu baseline failed: %s" % e)
        if args.batch_n > 0:
            try:
                _log("bench: batched instances, %d x %d steps" % (args.batch_n, args.batch_steps))
                extra["batch"] = batch_line(args.batch_n, args.batch_steps)
            except Exception as e:
                _log("bench: batch line failed: %s" % e)
            if not args.no_cpu and "batch" in extra:
                try:
                    _log("bench: cpu baseline of the batched workload")
                    extra["batch"]["cpu_baseline"] = batch_cpu_reference(extra["batch"]["sizes"])
                except Exception as e:
                    _log("bench: batch cpu baseline failed: %s" % e)
        if args.brio_wu:
            try:
                _log("bench: 1D Brio-Wu, 1024 cells to t = 0.1")
                extra["brio_wu"] = brio_wu_line(cpu=not args.no_cpu)
            except Exception as e:
                _log("bench: 1D line failed: %s" % e)
        if args.dsp_nx > 0:
            try:
                _log("bench: dissipative reconnection, %dx%d, %d steps" % (args.dsp_nx, args.dsp_ny, args.dsp_steps))
                extra["dissipative"] = dissipative_line(args.dsp_nx, args.dsp_ny, args.dsp_steps,
                                                        cpu=not args.no_cpu)
            except Exception as e:
                _log("bench: dissipative line failed: %s" % e)
        if args.grav_nx > 0:
            try:
                _log("bench: gravity (Rayleigh-Taylor), %dx%d, %d steps" % (args.grav_nx, args.grav_ny,
                                                                           args.grav_steps))
                extra["gravity"] = gravity_line(args.grav_nx, args.grav_ny, args.grav_steps, cpu=not args.no_cpu)
            except Exception as e:
                _log("bench: gravity line failed: %s" % e)
        if args.inverse_n > 0:
            try:
                _log("bench: velocity-mode inversion, %d instances x %d steps" % (args.inverse_n, args.inverse_steps))
                extra["inverse"] = inverse_line(args.inverse_n, args.inverse_steps)
            except Exception as e:
                _log("bench: inverse line failed: %s" % e)
        if args.adj_n > 0:
            try:
                _log("bench: adjoint, %d^2, %d steps" % (args.adj_n, args.adj_steps))
                extra["adjoint"] = adjoint_line(args.adj_n, args.adj_steps, rmn=args.adj_rmn)
            except Exception as e:
                _log("bench: adjoint line failed: %s" % e)

    line = {
        "metric": METRIC, "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.precision == 1 else "f64", "data": "synthetic",
        "config": {"workload": wname, "cells": list(p.interior), "solver": ["roe", "hlld", "lhlld", "mlau"][p.rmn],
                   "odr": p.odr, "r_k": p.r_k, "ctw": p.ctw, "cfl": p.cfl,
                   "precision": {0: "fp64", 1: "mixed", 2: "fp64_fma"}[args.precision],
                   "dt_final": dt_now, "finite": finite,
                   "launch": {None: "host (library default)", False: "host", True: "graph replay"}[capi.GRAPHS],
                   "launch_shape": shape},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    if not args.no_extra and p.ndim == 2:
        line["host"] = host_info()
        line["configs"] = config_table(line)
    print(json.dumps(line), flush=True)


def config_table(line):
    """GPU rate beside the CPU baseline for each of the five workloads, from
    the lines above (None where a line did not run)."""
    def g(d, *keys):
        for k in keys:
            if not isinstance(d, dict) or k not in d:
                return None
            d = d[k]
        return d
    cfg = {
        "1_brio_wu_1d_1024": {"gpu_fp64": g(line, "brio_wu", "value"), "cpu": g(line, "brio_wu", "cpu_baseline")},
        "2_turb2d_1024": {"gpu_fp64": line.get("value"), "cpu": line.get("cpu_baseline")},
        "3_khi_2048_mlau": {"gpu_fp64": g(line, "mixed", "fp64_value"), "gpu_mixed": g(line, "mixed", "mixed_value"),
                            "mixed_l1_rho": g(line, "mixed", "l1_rho"), "cpu": g(line, "mixed", "cpu_baseline")},
        "4_turb3d_512": {"gpu_fp64": g(line, "turb3d", "value"), "gpu_mixed": g(line, "turb3d", "mixed", "mixed_value"),
                         "mixed_l1_rho": g(line, "turb3d", "mixed", "l1_rho"), "cpu": g(line, "turb3d", "cpu_port")},
        "5_batch_64": {"gpu_fp64": g(line, "batch", "batched_value"), "cpu": g(line, "batch", "cpu_baseline"),
                       "adjoint_gradient_over_forward": g(line, "inverse", "ratio")},
    }
    for v in cfg.values():
        c = v.get("cpu")
        if isinstance(c, dict) and v.get("gpu_fp64") and c.get("value"):
            v["gpu_over_cpu"] = v["gpu_fp64"] / c["value"]
    return cfg


if __name__ == "__main__":
    main()
