"""Neural SDF surrogate on the GPU: the training kernels against the FP64
oracle (forward/backward gradients), the BF16 matrix-core inference kernel
against the oracle's BF16 arithmetic, and the surrogate inside the force
loop against the exact path."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from seismicmesh_amd import _lib
    _lib.require_gpu()


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))
    return t.to("cuda:0")


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_gemm_strided_and_split():
    import torch
    from seismicmesh_amd import _lib
    rng = np.random.default_rng(0)
    for (M, N, K, tA, tB, nsplit) in [(130, 70, 33, False, False, 1), (1, 256, 5000, True, False, 8),
                                      (256, 3, 777, True, True, 5), (300, 256, 256, False, True, 1),
                                      (257, 129, 1, False, False, 1)]:
        A = rng.standard_normal((K, M) if tA else (M, K)).astype(np.float32)
        B = rng.standard_normal((N, K) if tB else (K, N)).astype(np.float32)
        C0 = rng.standard_normal((M, N)).astype(np.float32)
        Ad, Bd, Cd = _dev(A), _dev(B), _dev(C0)
        sam, sak = (1, M) if tA else (K, 1)
        sbk, sbn = (1, K) if tB else (N, 1)
        work = torch.empty(nsplit * M * N + 1, dtype=torch.float32, device="cuda:0")
        _lib.check(_lib.lib.sm_gemm_f32(M, N, K, 0.5, _p(Ad), sam, sak, _p(Bd), sbk, sbn, 2.0,
                                        _p(Cd), N, nsplit, _p(work), None), "sm_gemm_f32")
        torch.cuda.synchronize()
        Am = A.T if tA else A
        Bm = B.T if tB else B
        ref = 0.5 * (Am.astype(np.float64) @ Bm.astype(np.float64)) + 2.0 * C0
        got = Cd.cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-5 * (np.abs(ref).max() + np.sqrt(K)), (M, N, K)


def _small_net(dim=3, seed=0):
    from seismicmesh_amd.neural import NeuralSDF
    return NeuralSDF(dim, [-1.0, 1.0] * dim, hidden=3, seed=seed)


def test_training_gradient_matches_oracle():
    from oracle import mlp as omlp
    for dim, lam, trunc in ((3, 0.1, 0.0), (2, 0.0, 0.0), (3, 0.1, 0.25)):
        net = _small_net(dim, seed=1)
        rng = np.random.default_rng(2)
        n = 300
        x = rng.uniform(-1, 1, (n, dim))
        t = (np.linalg.norm(x, axis=1) - 0.5).astype(np.float32)
        if trunc:
            t = np.clip(t, -trunc, trunc).astype(np.float32)
        part = net.loss_and_grad(_dev(x), _dev(t), eik=lam, eik_h=5e-3, trunc=trunc)
        g = net.grad.cpu().numpy().astype(np.float64)
        p = net.numpy_params().astype(np.float64)
        L, gref = omlp.loss_and_grad(p, dim, 256, 3, x, t.astype(np.float64), net.shift,
                                     net.scale, lam=lam, eik_h=5e-3, trunc=trunc)
        assert abs(float(part.cpu().numpy().astype(np.float64).sum()) - L) <= 1e-4 * abs(L)
        for name, (o, k) in net.layout.items():
            ref = gref[o:o + k]
            err = np.abs(g[o:o + k] - ref).max()
            # FP32 kernels against FP64: sums of up to 7 * 300 terms
            assert err <= 2e-3 * np.abs(ref).max() + 1e-7, (dim, name, err, np.abs(ref).max())


def test_adam_step_matches_formula():
    from seismicmesh_amd import _lib
    rng = np.random.default_rng(3)
    w, g = rng.standard_normal(1000).astype(np.float32), rng.standard_normal(1000).astype(np.float32)
    m1, m2 = rng.standard_normal(1000).astype(np.float32), rng.uniform(0, 1, 1000).astype(np.float32)
    wd, gd, m1d, m2d = _dev(w), _dev(g), _dev(m1), _dev(m2)
    _lib.check(_lib.lib.sm_mlp_adam(1000, _p(wd), _p(gd), _p(m1d), _p(m2d), 1e-3, 0.9, 0.999,
                                    1e-8, 7, None), "sm_mlp_adam")
    a = 0.9 * m1 + 0.1 * g
    b = 0.999 * m2 + 0.001 * g * g
    ref = w - 1e-3 * (a / (1 - 0.9 ** 7)) / (np.sqrt(b / (1 - 0.999 ** 7)) + 1e-8)
    assert np.allclose(wd.cpu().numpy(), ref, rtol=1e-5, atol=1e-7)


def test_inference_matches_bf16_arithmetic():
    """The MFMA kernel against the oracle's BF16 arithmetic (a wrong lane
    layout or k order would give O(1) errors); FP32 reference error is
    reported for scale."""
    from oracle import mlp as omlp
    for dim in (3, 2):
        net = _small_net(dim, seed=4)
        rng = np.random.default_rng(5)
        # non-trivial weights: scale the hidden layers so activations vary
        p = net.numpy_params()
        o, k = net.layout["w_hid"]
        p[o:o + k] *= 3.0
        net.set_params(p)
        x = rng.uniform(-1.2, 1.2, (5000, dim))
        got = net.eval(x).astype(np.float64)
        emu = omlp.forward_bf16(p.astype(np.float64), dim, 256, 3, x, net.shift, net.scale,
                                net.out_scale)
        P = omlp.unpack(p.astype(np.float64), dim, 256, 3)
        y32, _, _ = omlp.forward(P, (x - net.shift) * net.scale)
        ref32 = net.out_scale * y32
        span = np.abs(emu).max()
        err = np.abs(got - emu)
        print("dim %d: |bf16 kernel - bf16 oracle| max %.3e p99 %.3e median %.3e; "
              "max|kernel - fp64| %.3e; span %.3e"
              % (dim, err.max(), np.percentile(err, 99), np.median(err),
                 np.abs(got - ref32).max(), span))
        # FP32 accumulation (kernel) against FP64 sums (oracle) flips the BF16
        # rounding of an occasional activation (one BF16 ulp, 2^-8 relative);
        # the bulk of the points agree to FP32 round-off
        assert np.median(err) <= 1e-5 * span
        assert np.percentile(err, 99) <= 1e-3 * span
        assert err.max() <= 1e-2 * span


def test_centroid_inference_and_ghost_cells():
    net = _small_net(2, seed=6)
    rng = np.random.default_rng(7)
    p = rng.uniform(-1, 1, (50, 2))
    t = rng.integers(0, 50, (200, 3)).astype(np.int32)
    t[::7, 1] = -1
    got = net.eval_centroids(p, t)
    real = (t >= 0).all(axis=1)
    c = (p[t[real, 0]] + p[t[real, 1]] + p[t[real, 2]]) / 3.0
    assert np.array_equal(got[real], net.eval(c))
    assert np.isinf(got[~real]).all()


def test_training_learns_the_sdf():
    """Training with the library's kernels fits a ball's SDF (an answer
    that a correct forward/backward pass reaches in a few hundred steps)."""
    import seismicmesh_amd as sm
    from seismicmesh_amd.neural import NeuralSDF
    ball = sm.Ball([0.1, -0.2, 0.3], 0.5)
    net = NeuralSDF(3, [-1.0, 1.0] * 3, hidden=3, seed=0, omega=8.0)
    losses = net.train(ball, nsamples=100000, steps=600, batch=4096, lr=2e-3, lr_final=1e-4,
                       eik=0.1, near=0.05, seed=0, log_every=100)
    x = np.random.default_rng(11).uniform(-1, 1, (20000, 3))
    err = np.abs(net.eval(x) - ball.eval(x))
    print("ball surrogate: loss %s; |d_nn - d| mean %.4f p99 %.4f max %.4f"
          % ([(a, round(v, 6)) for a, v in losses], err.mean(), np.percentile(err, 99), err.max()))
    assert losses[-1][1] < 0.01 * losses[0][1]
    assert err.mean() < 0.005 and np.percentile(err, 99) < 0.03


def test_surrogate_loop_matches_exact_and_skips_work():
    """What the network gains: with the surrogate the force loop makes the
    same decisions (identical points) while most exact SDF evaluations of
    the centroid filter and the projection test are skipped.  The band is
    2 h or twice the network's largest error on a dense check of the box,
    whichever is larger."""
    import seismicmesh_amd as sm
    from seismicmesh_amd import generation
    from seismicmesh_amd.neural import NeuralSDF
    dom = sm.Difference([sm.Union([sm.Disk([-0.3, 0.0], 0.6),
                                   sm.Rectangle((0.0, 0.9, -0.4, 0.4))]),
                         sm.Disk([0.5, 0.1], 0.2)])
    net = NeuralSDF(2, (-1.0, 1.0, -1.0, 1.0), hidden=3, seed=0, omega=8.0)
    net.train(dom, nsamples=100000, steps=800, batch=4096, lr=2e-3, lr_final=1e-4, eik=0.1,
              near=0.05, seed=0)
    h = 0.02
    band, err = net.calibrate_band(dom, 2.0 * h, nval=200000)
    st = generation.prepare(dom, h, verbose=0, precision=32)
    a, b = generation.make_mesher(st), generation.make_mesher(st)
    b.set_surrogate(net, band)
    a.set_points(st["p"])
    b.set_points(st["p"])
    for _ in range(10):
        a.run(1)
        b.run(1)
        assert np.array_equal(a.points(), b.points())
    s = b.surrogate_stats()
    frac = s["skipped"] / (s["skipped"] + s["exact"])
    print("surrogate: max |d_nn - d| on the check %.4f, band %.4f (2h = %.3f): %d exact "
          "evaluations skipped, %d made (%.1f%% skipped), max |d_nn - d| in the band %.4f"
          % (err, band, 2 * h, s["skipped"], s["exact"], 100 * frac, s["max_err"]))
    assert frac > 0.3
    assert s["max_err"] < band
    a.close()
    b.close()


def test_surrogate_loop_3d_matches_exact():
    """3D: the surrogate-steered iterations (tetrahedra centroids and moved
    points) give the exact path's points on a small CSG."""
    import seismicmesh_amd as sm
    from seismicmesh_amd import generation
    from seismicmesh_amd.neural import NeuralSDF
    dom = sm.Difference([sm.Union([sm.Ball([-0.3, 0.0, 0.0], 0.55),
                                   sm.Cube((0.0, 0.8, -0.35, 0.35, -0.35, 0.35))]),
                         sm.Ball([0.45, 0.1, 0.0], 0.2)])
    net = NeuralSDF(3, (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0), hidden=3, seed=0, omega=8.0)
    # a truncated SDF: the band only needs the network near the surface
    net.train(dom, nsamples=200000, steps=2000, batch=4096, lr=2e-3, lr_final=1e-4, eik=0.1,
              near=0.05, seed=0, trunc=0.4)
    h = 0.08
    band, err = net.calibrate_band(dom, 2.0 * h, nval=100000)
    st = generation.prepare(dom, h, verbose=0, precision=32)
    a, b = generation.make_mesher(st), generation.make_mesher(st)
    b.set_surrogate(net, band)
    a.set_points(st["p"])
    b.set_points(st["p"])
    for _ in range(4):
        a.run(1)
        b.run(1)
        assert np.array_equal(a.points(), b.points())
    s = b.surrogate_stats()
    frac = s["skipped"] / (s["skipped"] + s["exact"])
    print("3D surrogate: max error %.4f, band %.4f: %.1f%% of exact evaluations skipped"
          % (err, band, 100 * frac))
    assert frac > 0.1
    a.close()
    b.close()
