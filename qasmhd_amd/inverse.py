"""Gradient-based inversion of initial velocity modes on many 2D domains.

Each instance is a periodic 2D decaying-MHD domain on [0, 2pi]^2 (the
turbulence2d set-up, rho = gam^2, p = gam so c_s = 1, rms|v| = rms|B| = M)
whose initial y velocity carries `nmodes` extra Fourier modes

    vy(a) = vy_base + sum_m a_m cos(kx_m x + ky_m y + phi_m),

0 < |k| <= kmax, wavevectors and phases drawn per instance from numpy
PCG64(seed * 1000 + i).  The parameters are the amplitudes a (nmodes per
instance).  With the dt schedule frozen (the schedule of the forward run at
the starting amplitudes a0, dt from setdt every second step as exec_ does,
2D/SERIAL/OTvortex/mhd2d_class.cpp:35-61), the loss is

    L(a) = sum_i sum_interior (rho_i(T; a_i) - rho_target_i)^2,

with rho_target_i the final density of the same run at amplitudes a_target
= a0 + a seeded perturbation.  dL/da comes from the hand-written discrete
adjoint of the step (qas_adj_forward / qas_adj_backward) chained through
the initial-condition map:

    my(a) = my_base + rho * D,   en(a) = en_base + D * (my_base + rho * D / 2),
    D = sum_m a_m psi_m,
    dL/da_m = sum_cells psi_m * (g_my * rho + g_en * my(a)).

The starting-amplitude forward that fixes the schedule runs all instances in
one batch (qas_batch_run: one launch per stage kernel for all domains, two
steps per graph replay).  The gradient runs batched too
(loss_and_grad_batched: qas_batch_adj_forward / qas_batch_adj_backward), or
per instance on host threads (loss_and_grad).  Losses at other amplitudes replay the frozen
schedule step by step (qas_step), so a central finite difference compares
two runs with identical step sizes.
"""
from __future__ import annotations

import time

import numpy as np

from . import capi, problems

NG = problems.NG


class VyModeInversion:
    def __init__(self, sizes, steps=200, seed=5, nmodes=8, kmax=4, mach=0.3, amp=0.05, perturb=0.02):
        self.sizes = [int(n) for n in sizes]
        self.steps = int(steps)
        self.nmodes = int(nmodes)
        self.base = []
        self.psi = []
        a0 = []
        at = []
        for i, n in enumerate(self.sizes):
            p = problems.turbulence2d(n=n, seed=seed * 1000 + i, vrms=mach, brms=mach)
            rng = np.random.Generator(np.random.PCG64(seed * 1000 + i))
            ks, _, ph = problems._modes2d(rng, nmodes, kmax)
            ny, nx = p.state.shape[1:]
            x = (np.arange(nx) - NG + 0.5) * p.dx
            y = (np.arange(ny) - NG + 0.5) * p.dy
            X, Y = np.meshgrid(x, y)
            self.psi.append(np.stack([np.cos(kx * X + ky * Y + f) for (kx, ky), f in zip(ks, ph)]))
            self.base.append(p)
            a = amp * rng.uniform(-1.0, 1.0, size=nmodes)
            a0.append(a)
            at.append(a + perturb * rng.standard_normal(nmodes))
        self.a0 = np.array(a0)
        self.a_target = np.array(at)
        self.dts = None
        self.rho_target = None

    @property
    def ninst(self):
        return len(self.sizes)

    def state(self, i, a):
        """Initial state (8, ny, nx) of instance i at amplitudes a."""
        p = self.base[i]
        st = p.state.copy()
        ro = st[0]
        D = np.tensordot(np.asarray(a, dtype=np.float64), self.psi[i], axes=1)
        my0 = st[2].copy()
        st[2] = my0 + ro * D
        st[7] = st[7] + D * (my0 + ro * D * 0.5)
        return st

    def _solver(self, i):
        return capi.solver_for(self.base[i])

    # ------------------------------------------------------------ forward runs
    def schedule(self, warm=False):
        """Batched forward at a0 (dt from setdt every 2nd step): fixes each
        instance's dt schedule; returns the final states at a0.  With `warm`,
        the run goes once first (launch-shape tuning, capture of the whole
        run as one graph) and the states are re-uploaded;
        `self.batch_seconds` is the wall time of the batched run."""
        solvers = [self._solver(i) for i in range(self.ninst)]
        try:
            for i, s in enumerate(solvers):
                s.upload(self.state(i, self.a0[i]))
            with capi.Batch(solvers) as b:
                if warm:
                    b.run(self.steps, flg=1, history=True)
                    for i, s in enumerate(solvers):
                        s.upload(self.state(i, self.a0[i]))
                        s.reset_counter()
                    for s in solvers:
                        s.sync()
                t0 = time.perf_counter()
                hist, _ = b.run(self.steps, flg=1, history=True)
                self.batch_seconds = time.perf_counter() - t0
            finals = [s.download() for s in solvers]
        finally:
            for s in solvers:
                s.close()
        self.dts = [np.array(h) for h in hist]
        return finals

    def final_state(self, i, a, solver=None):
        """Final state of instance i at amplitudes a over the frozen schedule."""
        if self.dts is None:
            self.schedule()
        s = solver or self._solver(i)
        try:
            s.upload(self.state(i, a))
            s.bound()
            for dt in self.dts[i]:
                s.step(dt)
            return s.download()
        finally:
            if solver is None:
                s.close()

    def targets(self):
        if self.rho_target is None:
            self.rho_target = [self.final_state(i, self.a_target[i])[0] for i in range(self.ninst)]
        return self.rho_target

    def instance_loss(self, i, a, solver=None):
        r = self.final_state(i, a, solver)[0] - self.targets()[i]
        return float(np.sum(r[NG:-NG, NG:-NG] ** 2))

    def loss(self, a):
        return sum(self.instance_loss(i, a[i]) for i in range(self.ninst))

    # ------------------------------------------------------------ gradient
    def instance_grad(self, i, a, solver=None):
        """(loss_i, dloss_i/da_i) by the discrete adjoint; `solver` (a
        context of instance i's shape, kept open) is used when given."""
        tgt = self.targets()[i]
        u0 = self.state(i, a)
        s = solver or self._solver(i)
        try:
            s.upload(u0)
            s.adj_forward(self.dts[i])
            uf = s.download()
            r = uf[0] - tgt
            g = np.zeros_like(uf)
            g[0, NG:-NG, NG:-NG] = 2.0 * r[NG:-NG, NG:-NG]
            gi = s.adj_backward(g)
        finally:
            if solver is None:
                s.close()
        w = gi[2] * u0[0] + gi[7] * u0[2]
        grad = np.tensordot(self.psi[i], w, axes=([1, 2], [0, 1]))
        return float(np.sum(r[NG:-NG, NG:-NG] ** 2)), grad

    def loss_and_grad(self, a, workers=4, keep=False):
        """Loss and gradient of all instances.  Each instance has its own
        context (its own HIP stream); `workers` host threads drive them, so
        up to that many instances' kernels run on the GPU side by side
        (small domains alone leave most of it idle).  The contexts stay open
        until every instance is done: freeing device memory waits for the
        whole device.  With `keep` they stay open for the next call (with
        their checkpoint room), until close().  Memory: each instance keeps
        8 x nd doubles per step of checkpoints (about 100 GB for the
        64-instance workload), so the contexts loss_and_grad_batched kept
        are freed first."""
        if self.dts is None:
            self.schedule()
        self.targets()
        self._free_batched()
        solvers = getattr(self, "_ipool", None) if keep else None
        if solvers is None:
            solvers = [self._solver(i) for i in range(self.ninst)]
            if keep:
                self._ipool = solvers
        try:
            if workers and workers > 1:
                from concurrent.futures import ThreadPoolExecutor
                with ThreadPoolExecutor(max_workers=workers) as ex:
                    res = list(ex.map(lambda i: self.instance_grad(i, a[i], solvers[i]), range(self.ninst)))
            else:
                res = [self.instance_grad(i, a[i], solvers[i]) for i in range(self.ninst)]
        finally:
            if not keep:
                for s in solvers:
                    s.close()
        L = 0.0
        G = np.zeros_like(np.asarray(a, dtype=np.float64))
        for i, (li, gi) in enumerate(res):
            L += li
            G[i] = gi
        return L, G

    def loss_and_grad_batched(self, a, groups=4, workers=4, backward="batched"):
        """Loss and gradient of all instances with the batched adjoint:
        the instances split into `groups` batches of about equal zone count,
        driven by `workers` host threads, each batch on its own stream, so
        that the batches' kernels run side by side (the adjoint line pass holds one
        wave per SIMD; other kernels fill the rest of each SIMD).  Per batch:
        the recorded forward of its instances in one batched run
        (qas_batch_adj_forward) and the backward with one launch per kernel
        for all of them (qas_batch_adj_backward), or, with backward
        "per_instance", each instance's own qas_adj_backward on the
        checkpoints the batched forward left in its context.  Same
        per-instance arithmetic as instance_grad.  The contexts, their
        checkpoint room and the batches are kept for the next call (an
        iterative inversion evaluates many gradients): close() frees them.
        Contexts loss_and_grad(keep=True) kept are freed first."""
        if backward not in ("batched", "per_instance"):
            raise ValueError("backward: 'batched' or 'per_instance'")
        if self.dts is None:
            self.schedule()
        tgt = self.targets()
        self._free_instances()
        groups = max(1, min(int(groups), self.ninst))
        pool = getattr(self, "_gpool", None)
        if pool is None or len(pool) != groups:
            self._free_batched()
            order = sorted(range(self.ninst), key=lambda i: -self.sizes[i] ** 2)
            members, load = [[] for _ in range(groups)], [0] * groups
            for i in order:  # largest first onto the lightest group
                k = load.index(min(load))
                members[k].append(i)
                load[k] += self.sizes[i] ** 2
            pool = []
            for mem in members:
                mem = sorted(mem)
                solvers = [self._solver(i) for i in mem]
                pool.append((mem, solvers, capi.Batch(solvers)))
            self._gpool = pool
        u0 = [self.state(i, a[i]) for i in range(self.ninst)]
        gi = [None] * self.ninst
        Ls = [0.0] * self.ninst

        def run(group):
            mem, solvers, b = group
            for i, s in zip(mem, solvers):
                s.upload(u0[i])
                s.reset_counter()
            b.adj_forward(np.array([self.dts[i] for i in mem]))
            gfin = []
            for i, s in zip(mem, solvers):
                r = s.download()[0] - tgt[i]
                g = np.zeros((8,) + s.shape)
                g[0, NG:-NG, NG:-NG] = 2.0 * r[NG:-NG, NG:-NG]
                gfin.append(g)
                Ls[i] = float(np.sum(r[NG:-NG, NG:-NG] ** 2))
            if backward == "per_instance":
                for i, s, g in zip(mem, solvers, gfin):
                    gi[i] = s.adj_backward(g)
            else:
                for i, g in zip(mem, b.adj_backward(gfin)):
                    gi[i] = g

        if len(pool) > 1 and workers > 1:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(max_workers=min(len(pool), workers)) as ex:
                list(ex.map(run, pool))
        else:
            run(pool[0])
        L = sum(Ls)
        G = np.zeros_like(np.asarray(a, dtype=np.float64))
        for i in range(self.ninst):
            w = gi[i][2] * u0[i][0] + gi[i][7] * u0[i][2]
            G[i] = np.tensordot(self.psi[i], w, axes=([1, 2], [0, 1]))
        return L, G

    def _free_batched(self):
        pool = getattr(self, "_gpool", None)
        self._gpool = None
        for _, solvers, b in pool or []:
            b.close()
            for s in solvers:
                s.close()

    def _free_instances(self):
        ipool = getattr(self, "_ipool", None)
        self._ipool = None
        for s in ipool or []:
            s.close()

    def close(self):
        """Free the contexts loss_and_grad_batched and loss_and_grad(keep)
        keep (at most one of the two sets is held at a time)."""
        self._free_batched()
        self._free_instances()

    def fd_grad(self, i, a, m, h=1e-6, richardson=True):
        """Central difference of instance i's loss in a_m over the frozen
        schedule (Richardson-extrapolated from steps h and h/2)."""
        with self._solver(i) as s:
            def D(hh):
                ap = np.array(a, dtype=np.float64)
                am = ap.copy()
                ap[m] += hh
                am[m] -= hh
                return (self.instance_loss(i, ap, s) - self.instance_loss(i, am, s)) / (2.0 * hh)
            if not richardson:
                return D(h)
            return (4.0 * D(h / 2) - D(h)) / 3.0


def config_sizes(ninst=64, seed=5, choices=(128, 192, 256, 384, 512)):
    """Square domain sizes drawn uniformly (numpy PCG64(seed)) from `choices`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [int(v) for v in rng.choice(list(choices), size=ninst)]
