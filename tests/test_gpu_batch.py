"""Many domains in one handle (GPU) against one handle per domain.

Each domain of a batch has its own triangulation -- of its convex hull,
or inside its own far square -- so the batch's cells and edges of every
domain are identical to a single-domain handle's triangulating the same
region on the same points in general position; point updates differ only
through the order of the FP32 L0 sums
(per-domain reduction instead of the single handle's block partials), so
one iteration agrees to 1e-5 of the update."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from seismicmesh_amd import _lib
    _lib.require_gpu()


def _domains():
    import seismicmesh_amd as sm
    return ([sm.Disk([0.0, 0.0], 1.0),
             sm.Rectangle((-1.0, 0.5, -0.5, 1.0)),
             sm.Difference([sm.Rectangle((-1.0, 1.0, -1.0, 1.0)), sm.Disk([0.3, 0.2], 0.4)]),
             sm.Union([sm.Disk([-0.5, 0.0], 0.6), sm.Rectangle((0.0, 1.2, -0.3, 0.3))])],
            [0.09, 0.07, 0.08, 0.06])


def _jitter(sts, seed=7):
    # the initial lattices are degenerate (cocircular quadruples), where the
    # Delaunay triangulation is not unique and insertion order (which the
    # far vertices change) breaks ties: perturb the free points by 1e-6 h so
    # the comparison is between unique answers
    rng = np.random.default_rng(seed)
    for st in sts:
        p = st["p"].copy()
        p[st["nfix"]:] += rng.uniform(-1e-6, 1e-6, p[st["nfix"]:].shape) * st["h0"]
        st["p"] = p
    return sts


@pytest.mark.parametrize("far_square", [False, True], ids=["hull", "far_square"])
def test_batch_matches_single_handles(far_square):
    from seismicmesh_amd import batch, generation
    doms, hs = _domains()
    sts = _jitter(batch.prepare_batch(doms, hs))
    bm = batch.make_batch_mesher(sts, far_square=far_square)
    bm.set_graph(False)
    bm.set_points(np.vstack([st["p"] for st in sts]))
    bm.triangulate()
    parts = bm.split(bm.points(), bm.cells())
    be = bm.edges()
    singles = []
    for d, st in enumerate(sts):
        m = generation.make_mesher(st, far_square=far_square)
        m.set_graph(False)
        m.set_points(st["p"])
        m.triangulate()
        got = {tuple(sorted(c)) for c in parts[d][1].tolist()}
        assert got == {tuple(sorted(c)) for c in m.cells().tolist()}, "domain %d" % d
        b, e = bm.voff[d], bm.voff[d + 1]
        mine = be[(be[:, 0] >= b) & (be[:, 0] < e)] - b
        assert np.array_equal(mine, m.edges()), "domain %d" % d
        singles.append(m)
    bm.run(1)
    pb = bm.points()
    for d, (st, m) in enumerate(zip(sts, singles)):
        m.run(1)
        b, e = bm.voff[d], bm.voff[d + 1]
        dp_single = m.points() - st["p"]
        dp_batch = pb[b:e] - st["p"]
        scale = np.abs(dp_single).max()
        assert scale > 0
        assert np.abs(dp_batch - dp_single).max() <= 1e-5 * scale, "domain %d" % d
        m.close()
    bm.close()


@pytest.mark.parametrize("far_square", [False, True], ids=["hull", "far_square"])
def test_batch_graph_matches_plain(far_square):
    from seismicmesh_amd import batch
    doms, hs = _domains()
    sts = batch.prepare_batch(doms, hs)
    a = batch.make_batch_mesher(sts, far_square=far_square)
    b = batch.make_batch_mesher(sts, far_square=far_square)
    a.set_graph(False)
    b.set_graph(True)
    p0 = np.vstack([st["p"] for st in sts])
    a.set_points(p0)
    b.set_points(p0)
    for _ in range(12):
        a.run(1)
        b.run(1)
        assert np.array_equal(a.points(), b.points())
    a.close()
    b.close()


def test_generate_mesh_batch_against_single_runs():
    import seismicmesh_amd as sm
    from seismicmesh_amd import batch, mesh_utils as mu
    doms, hs = _domains()
    out = batch.generate_mesh_batch(doms, hs)
    for d, ((p, t), dom, h) in enumerate(zip(out, doms, hs)):
        q, s = sm.generate_mesh(dom, h, verbose=0, precision=32)
        assert abs(len(p) - len(q)) <= 0.02 * len(q) + 2, "domain %d" % d
        qual_b, qual_s = mu.simp_qual(p, t), mu.simp_qual(q, s)
        assert qual_b.min() > 0.3 and abs(qual_b.mean() - qual_s.mean()) < 0.01, "domain %d" % d


def test_many_domains_workload_small():
    from seismicmesh_amd import batch, workloads
    doms, hs = workloads.many_domains(n=8, nmax=4000)
    out = batch.generate_mesh_batch(doms, hs, max_iter=30)
    from seismicmesh_amd import mesh_utils as mu
    for p, t in out:
        assert len(t) > 0 and mu.simp_qual(p, t).min() > 0.1
