"""Batched domains (qas_batch_*): every instance of a batch of uneven 2D
grids ends bit-identical to its own single-domain qas_run -- state and dt
history -- which in turn equals the reference (test_gpu_parity.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems():
    from qasmhd_amd import problems
    sizes = [(32, 32), (48, 24), (24, 56), (64, 40)]
    return [problems.turbulence2d(n=nx, ny=ny, seed=10 + i, vrms=0.3 + 0.2 * i) for i, (nx, ny) in enumerate(sizes)]


@pytest.mark.parametrize("precision,nsteps", [(0, 7), (0, 10), (1, 6)])
def test_batch_matches_single_runs(precision, nsteps):
    from qasmhd_amd import capi
    print("case: batch precision", precision, "steps", nsteps, flush=True)
    ps = _problems()
    ref = []
    for p in ps:
        with capi.solver_for(p, precision=precision) as s:
            s.upload(p.state)
            h, dt = s.run(nsteps, flg=1)
            ref.append((h, dt, s.download()))
    solvers = [capi.solver_for(p, precision=precision) for p in ps]
    for s, p in zip(solvers, ps):
        s.upload(p.state)
    with capi.Batch(solvers) as b:
        hist, dts = b.run(nsteps, flg=1)
        # a second call continues the step counter as qas_run does
        hist2, dts2 = b.run(3, flg=1)
    outs = [s.download() for s in solvers]
    for i, p in enumerate(ps):
        assert np.array_equal(hist[i], ref[i][0]), i
        assert dts[i] == ref[i][1], i
    # compare the continued run with single-domain runs of nsteps + 3
    for i, p in enumerate(ps):
        with capi.solver_for(p, precision=precision) as s:
            s.upload(p.state)
            s.run(nsteps, flg=1)
            h3, d3 = s.run(3, flg=1)
            full = s.download()
        assert np.array_equal(hist2[i], h3), i
        assert dts2[i] == d3, i
        assert np.array_equal(outs[i], full), i
    for s in solvers:
        s.close()


@pytest.mark.parametrize("shape", [(0, 4, 256), (1, 4, 256), (0, 2, 128), (1, 1, 64), (1, 2, 256)])
def test_batch_shapes_are_bit_identical(shape):
    # every launch shape the tuner may pick (block order, x-sweep wavefronts
    # per workgroup, element-wise workgroup size) gives the single-run result
    from qasmhd_amd import capi
    print("case: batch shape", shape, flush=True)
    ps = _problems()
    ref = []
    for p in ps:
        with capi.solver_for(p) as s:
            s.upload(p.state)
            h, dt = s.run(6, flg=1)
            ref.append((h, dt, s.download()))
    solvers = [capi.solver_for(p) for p in ps]
    for s, p in zip(solvers, ps):
        s.upload(p.state)
    with capi.Batch(solvers) as b:
        b.set_shape(*shape)
        hist, dts = b.run(6, flg=1)
        assert b.shape() == shape
    for i, s in enumerate(solvers):
        assert np.array_equal(hist[i], ref[i][0]), i
        assert dts[i] == ref[i][1], i
        assert np.array_equal(s.download(), ref[i][2]), i
        s.close()


def test_batch_host_launches_equal_graph_replay():
    # qas_batch_use_graphs(b, 0): every kernel launched from the host gives
    # the graph replay's states and dt histories (and the single runs')
    from qasmhd_amd import capi
    print("case: batch host launches vs graph replay", flush=True)
    ps = _problems()
    outs = {}
    for graphs in (True, False):
        solvers = [capi.solver_for(p) for p in ps]
        for s, p in zip(solvers, ps):
            s.upload(p.state)
        with capi.Batch(solvers) as b:
            b.use_graphs(graphs)
            hist, dts = b.run(7, flg=1)
            hist2, _ = b.run(4, flg=1)
        outs[graphs] = (hist, dts, hist2, [s.download() for s in solvers])
        for s in solvers:
            s.close()
    a, h = outs[True], outs[False]
    assert np.array_equal(a[0], h[0]) and np.array_equal(a[1], h[1]) and np.array_equal(a[2], h[2])
    for i in range(len(ps)):
        assert np.array_equal(a[3][i], h[3][i]), i


def test_batch_run_longer_than_one_graph():
    # a run is captured as one graph of up to 128 step pairs: 515 steps are
    # two launches of that graph and three host-launched steps, and give
    # every domain its own qas_run's state and dt history
    from qasmhd_amd import capi
    print("case: batch run of 515 steps (graph cap)", flush=True)
    ps = _problems()[:3]
    ref = []
    for p in ps:
        with capi.solver_for(p) as s:
            s.upload(p.state)
            h, dt = s.run(515, flg=1)
            ref.append((h, dt, s.download()))
    solvers = [capi.solver_for(p) for p in ps]
    for s, p in zip(solvers, ps):
        s.upload(p.state)
    with capi.Batch(solvers) as b:
        hist, dts = b.run(515, flg=1)
    for i, s in enumerate(solvers):
        assert np.array_equal(hist[i], ref[i][0]), i
        assert dts[i] == ref[i][1], i
        assert np.array_equal(s.download(), ref[i][2]), i
        s.close()


def test_batch_autotune_picks_a_candidate():
    from qasmhd_amd import capi
    print("case: batch autotune", flush=True)
    ps = _problems()
    solvers = [capi.solver_for(p) for p in ps]
    for s, p in zip(solvers, ps):
        s.upload(p.state)
    with capi.Batch(solvers) as b:
        order, xwpb, ewb = b.tune()
        assert order in (0, 1) and xwpb in (2, 4) and ewb in (128, 256)
        hist, _ = b.run(2, flg=1)
        assert b.shape() == (order, xwpb, ewb)
    for s in solvers:
        s.close()


def test_batch_rejects_mixed_schemes():
    from qasmhd_amd import capi, problems
    p = problems.turbulence2d(n=16, seed=1)
    a = capi.solver_for(p)
    b = capi.solver_for(p, rmn=3)
    a.upload(p.state)
    b.upload(p.state)
    with pytest.raises(capi.QasError):
        capi.Batch([a, b])
    a.close()
    b.close()


def test_batch_history_growth_then_single_run():
    """A batch run with history over more steps than a context's history
    buffer reallocates that buffer; the context's own step-pair graph
    (captured by an earlier qas_run with history, graph replay on) must not
    keep writing the freed one.  s.run(4) -> batch.run(8) -> s.run(4): the
    last history and state equal those of a fresh context run over the same
    16 steps."""
    from qasmhd_amd import capi
    print("case: batch history growth", flush=True)
    ps = _problems()[:2]
    solvers = [capi.solver_for(p) for p in ps]
    for s, p in zip(solvers, ps):
        s.lib.qas_use_graphs(s.ctx, 1)
        s.upload(p.state)
        s.run(4, flg=1)
    with capi.Batch(solvers) as b:
        b.run(8, flg=1)
    last = [s.run(4, flg=1) for s in solvers]
    outs = [s.download() for s in solvers]
    for s in solvers:
        s.close()
    for i, p in enumerate(ps):
        with capi.solver_for(p) as s:
            s.upload(p.state)
            s.run(12, flg=1)
            h, dt = s.run(4, flg=1)
            ref = s.download()
        assert np.array_equal(last[i][0], h), i
        assert last[i][1] == dt, i
        assert np.array_equal(outs[i], ref), i
