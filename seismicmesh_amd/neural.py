"""Neural signed-distance surrogate of a domain (no counterpart in the
reference, which evaluates its CSG functions directly,
SeismicMesh/geometry/signed_distance_functions.py:283-623).

NeuralSDF is a multilayer perceptron dim -> 256 -> ... -> 256 -> 1 with
SiLU activations.  It is trained on the GPU by libsmgpu's own FP32 kernels
(sm_gemm_f32 for the forward and backward products, sm_mlp_bias_silu,
sm_mlp_silu_backward, sm_mlp_sdf_loss, sm_mlp_adam) on samples of the
domain's exact SDF -- half uniform in the bounding box, half within `near`
of the surface -- with the squared error plus a central-difference eikonal
term.  Inference (sm_mlp_eval) is the fused BF16 matrix-core kernel.

The force loop uses it through Mesher.set_surrogate: the network answers
the centroid filter and the projection test away from the boundary; inside
the band |d_nn| < band the exact program answers, and it records the
network's largest error there.  PyTorch only provides device buffers.
"""
import ctypes

import numpy as np

from . import _lib

__all__ = ["NeuralSDF", "MLPStruct"]


class MLPStruct(ctypes.Structure):
    """sm_mlp (smgpu.h)."""
    _fields_ = [
        ("in_dim", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("hidden", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("w_in", ctypes.c_void_p),
        ("b_in", ctypes.c_void_p),
        ("w_hid", ctypes.c_void_p),
        ("b_hid", ctypes.c_void_p),
        ("w_out", ctypes.c_void_p),
        ("b_out", ctypes.c_float),
        ("reserved2", ctypes.c_float),
        ("shift", ctypes.c_double * 3),
        ("scale", ctypes.c_double * 3),
        ("out_scale", ctypes.c_double),
    ]


def param_layout(dim, width, hidden):
    """(name, offset, size) of each parameter block in the flat FP32 vector:
    w_in (width, dim), b_in (width), w_hid (hidden, width, width) with rows
    = output features, b_hid (hidden, width), w_out (width), b_out (1)."""
    sizes = [("w_in", width * dim), ("b_in", width), ("w_hid", hidden * width * width),
             ("b_hid", hidden * width), ("w_out", width), ("b_out", 1)]
    out, off = {}, 0
    for name, n in sizes:
        out[name] = (off, n)
        off += n
    return out, off


def init_params(dim, width, hidden, seed=0, omega=1.0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases; the
    input layer's bound is multiplied by `omega` (its ridge functions then
    vary on a length scale ~1/omega of the box, which lets the network
    resolve the creases and small parts of a CSG sooner)."""
    lay, total = param_layout(dim, width, hidden)
    rng = np.random.default_rng(seed)
    p = np.empty(total, dtype=np.float32)
    fan = {"w_in": dim, "b_in": dim, "w_hid": width, "b_hid": width, "w_out": width,
           "b_out": width}
    for name, (off, n) in lay.items():
        b = 1.0 / np.sqrt(fan[name])
        if name in ("w_in", "b_in"):
            b *= omega
        p[off:off + n] = rng.uniform(-b, b, n)
    return p


def _ptr(t, offset_elems=0):
    return ctypes.c_void_p(t.data_ptr() + offset_elems * t.element_size())


class NeuralSDF:
    """A trainable surrogate of a 2D or 3D signed distance function on the
    box `bbox` (x0, x1, y0, y1[, z0, z1]).  Inputs are mapped to
    x' = (x - centre) * s with s = 2 / (largest side), and the network
    predicts d * s, so |grad| = 1 is the eikonal condition in both frames."""

    WIDTH = 256

    def __init__(self, dim, bbox, hidden=3, seed=0, device=0, omega=1.0):
        import torch

        _lib.require_gpu()
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.dim, self.hidden, self.width = dim, int(hidden), self.WIDTH
        b = np.asarray(bbox, dtype=np.float64).reshape(dim, 2)
        self.shift = (b[:, 0] + b[:, 1]) / 2.0
        s = 2.0 / float((b[:, 1] - b[:, 0]).max())
        self.scale = np.full(dim, s)
        self.out_scale = 1.0 / s
        self.bbox = b
        self.layout, self.nparams = param_layout(dim, self.width, self.hidden)
        self.params = torch.from_numpy(init_params(dim, self.width, self.hidden, seed, omega)).to(self.dev)
        self.grad = torch.zeros_like(self.params)
        self.m1 = torch.zeros_like(self.params)
        self.m2 = torch.zeros_like(self.params)
        self.w_hid_bf16 = torch.empty(max(self.hidden * self.width * self.width, 8),
                                      dtype=torch.bfloat16, device=self.dev)
        self.step = 0
        self.trunc = None
        self._bufs = {}
        self._dirty = True
        self._shift_d = torch.tensor(np.r_[self.shift, 0.0][:3], dtype=torch.float64,
                                     device=self.dev)
        self._scale_d = torch.tensor(np.r_[self.scale, 0.0][:3], dtype=torch.float64,
                                     device=self.dev)

    # ------------------------------------------------------------ helpers
    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def _p(self, name, buf=None):
        buf = self.params if buf is None else buf
        return _ptr(buf, self.layout[name][0])

    def numpy_params(self):
        return self.params.cpu().numpy().copy()

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.shape != (self.nparams,):
            raise ValueError("expected %d parameters" % self.nparams)
        self.params.copy_(self.torch.from_numpy(flat))
        self._dirty = True

    def struct(self):
        """The sm_mlp view of the current weights (hidden weights rounded to
        BF16 and packed in fragment order on the device when they changed)."""
        if self._dirty:
            if self.hidden:
                _lib.check(_lib.lib.sm_mlp_pack_hidden(self.hidden, self.width, self._p("w_hid"),
                                                       _ptr(self.w_hid_bf16), self._stream()),
                           "sm_mlp_pack_hidden")
            self._dirty = False
        st = MLPStruct()
        st.in_dim, st.width, st.hidden = self.dim, self.width, self.hidden
        st.w_in = self._p("w_in").value
        st.b_in = self._p("b_in").value
        st.w_hid = self.w_hid_bf16.data_ptr()
        st.b_hid = self._p("b_hid").value
        st.w_out = self._p("w_out").value
        st.b_out = float(self.params[self.layout["b_out"][0]].item())
        for i in range(3):
            st.shift[i] = float(self.shift[i]) if i < self.dim else 0.0
            st.scale[i] = float(self.scale[i]) if i < self.dim else 0.0
        st.out_scale = self.out_scale
        self._struct = st
        return st

    # ---------------------------------------------------------- inference
    def eval_device(self, xt):
        """d_nn at device points xt ((n, dim) float64 tensor) -> float32 tensor."""
        n = xt.shape[0]
        out = self.torch.empty(n, dtype=self.torch.float32, device=self.dev)
        st = self.struct()
        _lib.check(_lib.lib.sm_mlp_eval(ctypes.byref(st), _ptr(xt), n, _ptr(out), self._stream()),
                   "sm_mlp_eval")
        return out

    def eval(self, x):
        """d_nn at host points x (n, dim) -> float32 numpy array."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, self.dim))
        if len(x) == 0:
            return np.zeros(0, dtype=np.float32)
        xt = self.torch.from_numpy(x).to(self.dev)
        out = self.eval_device(xt)
        return out.cpu().numpy()

    def eval_centroids(self, p, t):
        """d_nn at the centroids of cells t of points p (cells with a
        negative id give +inf)."""
        pt = self.torch.from_numpy(np.ascontiguousarray(p, dtype=np.float64)).to(self.dev)
        tt = self.torch.from_numpy(np.ascontiguousarray(t, dtype=np.int32)).to(self.dev)
        out = self.torch.empty(len(t), dtype=self.torch.float32, device=self.dev)
        st = self.struct()
        _lib.check(_lib.lib.sm_mlp_eval_centroids(ctypes.byref(st), _ptr(tt), len(t), _ptr(pt),
                                                  _ptr(out), self._stream()),
                   "sm_mlp_eval_centroids")
        return out.cpu().numpy()

    # ----------------------------------------------------------- training
    def _buffers(self, n, eik):
        rows = n * (2 * self.dim + 1 if eik else 1)
        key = (n, bool(eik))
        if key in self._bufs:
            return self._bufs[key]
        torch, W, L = self.torch, self.width, self.hidden
        f32 = dict(dtype=torch.float32, device=self.dev)
        nsplit = int(max(1, min(64, rows // 2048)))
        b = dict(rows=rows, nsplit=nsplit,
                 X=torch.empty(rows * self.dim, **f32),
                 Z=[torch.empty(rows * W, **f32) for _ in range(L + 1)],
                 A=[torch.empty(rows * W, **f32) for _ in range(L + 1)],
                 dA=torch.empty(rows * W, **f32), dZ=torch.empty(rows * W, **f32),
                 y=torch.empty(rows, **f32), gy=torch.empty(rows, **f32),
                 ones=torch.from_numpy(np.ones(rows, dtype=np.float32)).to(self.dev),
                 part=torch.empty((n + 255) // 256, **f32),
                 work=torch.empty(nsplit * W * W + 16, **f32))
        self._bufs = {key: b}  # keep one set
        return b

    def _gemm(self, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, nsplit=1, work=None,
              alpha=1.0, beta=0.0):
        _lib.check(_lib.lib.sm_gemm_f32(M, N, K, alpha, A, sam, sak, B, sbk, sbn, beta, C, ldc,
                                        nsplit, work if work is not None else None,
                                        self._stream()), "sm_gemm_f32")

    def loss_and_grad(self, xt, tt, eik=0.1, eik_h=5e-3, trunc=0.0):
        """Forward and backward pass over the samples xt ((n, dim) float64
        device tensor) with targets tt (float32, distances in the network's
        frame, i.e. d * s; truncated at +-trunc when trunc > 0, where the
        eikonal term is dropped).  Fills self.grad; returns the loss
        partials (device)."""
        n = xt.shape[0]
        W, L, dim = self.width, self.hidden, self.dim
        b = self._buffers(n, eik > 0)
        rows, nsplit, work = b["rows"], b["nsplit"], _ptr(b["work"])
        st = self._stream()
        _lib.check(_lib.lib.sm_mlp_make_inputs(n, dim, _ptr(xt), _ptr(self._shift_d),
                                               _ptr(self._scale_d),
                                               float(eik_h) if eik > 0 else 0.0,
                                               _ptr(b["X"]), st), "sm_mlp_make_inputs")
        Z, A = b["Z"], b["A"]
        # forward: Z0 = X W_in^T, A_l = silu(Z_l + b_l), y = A_L w_out + b_out
        self._gemm(rows, W, dim, _ptr(b["X"]), dim, 1, self._p("w_in"), 1, dim, _ptr(Z[0]), W)
        _lib.check(_lib.lib.sm_mlp_bias_silu(rows, W, _ptr(Z[0]), self._p("b_in"), _ptr(A[0]),
                                             st), "sm_mlp_bias_silu")
        woff = self.layout["w_hid"][0]
        boff = self.layout["b_hid"][0]
        for l in range(1, L + 1):
            Wl = _ptr(self.params, woff + (l - 1) * W * W)
            self._gemm(rows, W, W, _ptr(A[l - 1]), W, 1, Wl, 1, W, _ptr(Z[l]), W)
            _lib.check(_lib.lib.sm_mlp_bias_silu(rows, W, _ptr(Z[l]),
                                                 _ptr(self.params, boff + (l - 1) * W),
                                                 _ptr(A[l]), st), "sm_mlp_bias_silu")
        self._gemm(rows, 1, W, _ptr(A[L]), W, 1, self._p("w_out"), 1, 0, _ptr(b["y"]), 1)
        _lib.check(_lib.lib.sm_mlp_bias_silu(rows, 1, _ptr(b["y"]), self._p("b_out"), None, st),
                   "sm_mlp_bias_silu")
        _lib.check(_lib.lib.sm_mlp_sdf_loss(n, dim, _ptr(b["y"]), _ptr(tt), float(eik_h),
                                            float(eik), float(trunc), _ptr(b["gy"]),
                                            _ptr(b["part"]), st),
                   "sm_mlp_sdf_loss")
        g = self.grad
        gy = _ptr(b["gy"])
        # output layer: dw_out = gy^T A_L, db_out = sum gy, dA_L = gy w_out^T
        self._gemm(1, W, rows, gy, 0, 1, _ptr(A[L]), W, 1, self._p("w_out", g), W, nsplit, work)
        self._gemm(1, 1, rows, gy, 0, 1, _ptr(b["ones"]), 1, 0, self._p("b_out", g), 1, nsplit,
                   work)
        self._gemm(rows, W, 1, gy, 1, 0, self._p("w_out"), 0, 1, _ptr(b["dA"]), W)
        for l in range(L, -1, -1):
            _lib.check(_lib.lib.sm_mlp_silu_backward(rows * W, _ptr(Z[l]), _ptr(b["dA"]),
                                                     _ptr(b["dZ"]), st), "sm_mlp_silu_backward")
            dZ = _ptr(b["dZ"])
            if l == 0:
                self._gemm(W, dim, rows, dZ, 1, W, _ptr(b["X"]), dim, 1, self._p("w_in", g), dim,
                           nsplit, work)
                self._gemm(1, W, rows, _ptr(b["ones"]), 0, 1, dZ, W, 1, self._p("b_in", g), W,
                           nsplit, work)
                break
            # dW_l = dZ_l^T A_{l-1}, db_l = sum_rows dZ_l, dA_{l-1} = dZ_l W_l
            self._gemm(W, W, rows, dZ, 1, W, _ptr(A[l - 1]), W, 1,
                       _ptr(g, woff + (l - 1) * W * W), W, nsplit, work)
            self._gemm(1, W, rows, _ptr(b["ones"]), 0, 1, dZ, W, 1, _ptr(g, boff + (l - 1) * W),
                       W, nsplit, work)
            self._gemm(rows, W, W, dZ, W, 1, _ptr(self.params, woff + (l - 1) * W * W), W, 1,
                       _ptr(b["dA"]), W)
        return b["part"]

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(_lib.lib.sm_mlp_adam(self.nparams, _ptr(self.params), _ptr(self.grad),
                                        _ptr(self.m1), _ptr(self.m2), float(lr), float(b1),
                                        float(b2), float(eps), self.step, self._stream()),
                   "sm_mlp_adam")
        self._dirty = True

    @staticmethod
    def sample(domain, bbox, n, near, seed=0, pad=0.05):
        """n points: n/2 uniform in the padded box, n/2 with |d| < near
        (rejection from uniform candidates); returns (x, exact d)."""
        rng = np.random.default_rng(seed)
        b = np.asarray(bbox, dtype=np.float64)
        lo = b[:, 0] - pad * (b[:, 1] - b[:, 0])
        hi = b[:, 1] + pad * (b[:, 1] - b[:, 0])
        dim = len(b)
        nu = n // 2
        xu = rng.uniform(lo, hi, (nu, dim))
        near_x = []
        need = n - nu
        while need > 0:
            cand = rng.uniform(lo, hi, (max(4 * need, 65536), dim))
            dc = domain.eval(cand)
            sel = cand[np.abs(dc) < near]
            near_x.append(sel[:need])
            need -= len(near_x[-1])
        x = np.vstack([xu] + near_x)
        return x, domain.eval(x)

    def train(self, domain, nsamples=200000, steps=1500, batch=8192, lr=1e-3, lr_final=5e-5,
              eik=0.1, eik_h=5e-3, near=None, seed=0, log_every=0, samples=None, trunc=None):
        """Fit the exact SDF of `domain` (truncated at +-trunc when given:
        the network then only has to be accurate within trunc of the
        surface, which is all the exact-fallback band needs); returns the
        loss every `log_every` steps (all on the GPU; the loss is read back
        only when logged)."""
        torch = self.torch
        if near is None:
            near = 0.05 * float((self.bbox[:, 1] - self.bbox[:, 0]).max()) / 2.0
        if samples is None:
            x, d = self.sample(domain, self.bbox, nsamples, near, seed)
        else:
            x, d = samples
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.trunc = float(trunc) if trunc else None
        d = np.asarray(d, dtype=np.float64)
        if self.trunc:
            d = np.clip(d, -self.trunc, self.trunc)
        tn = (d * self.scale[0]).astype(np.float32)
        trunc_n = self.trunc * self.scale[0] if self.trunc else 0.0
        n = len(x)
        batch = min(batch, n)
        rng = np.random.default_rng(seed + 1)
        pos = n
        losses = []
        for it in range(steps):
            if pos + batch > n:
                # a new epoch: shuffle on the host, upload
                perm = rng.permutation(n)
                xs = torch.from_numpy(np.ascontiguousarray(x[perm])).to(self.dev)
                ts = torch.from_numpy(tn[perm]).to(self.dev)
                pos = 0
            xb, tb = xs[pos:pos + batch], ts[pos:pos + batch]
            pos += batch
            part = self.loss_and_grad(xb, tb, eik=eik, eik_h=eik_h, trunc=trunc_n)
            frac = it / max(steps - 1, 1)
            rate = lr * (lr_final / lr) ** frac
            self.adam_step(rate)
            if log_every and (it % log_every == 0 or it == steps - 1):
                losses.append((it, float(part.cpu().numpy().astype(np.float64).sum())))
        torch.cuda.synchronize(self.dev)
        return losses

    def calibrate_band(self, domain, band, nval=100000, safety=2.0, seed=12345):
        """The exact-fallback band to use with this network: at least `band`
        (e.g. 2 h) and at least `safety` times the largest |d_nn - d| over
        nval uniform points of the box (judged by the exact program; both
        truncated for a network trained on a truncated SDF, whose band must
        stay below the truncation).  Returns (band, max error)."""
        rng = np.random.default_rng(seed)
        x = rng.uniform(self.bbox[:, 0], self.bbox[:, 1], (nval, self.dim))
        err = self.max_error(domain, x)
        out = max(float(band), safety * err)
        if self.trunc and out >= self.trunc:
            raise ValueError("the band %.4g reaches the truncation %.4g of the training "
                             "targets: train longer or truncate wider" % (out, self.trunc))
        return out, err

    def calibrate_bands(self, domain, band, res=64, nval=8000000, near=None, safety=2.0,
                        seed=54321):
        """Per-region bands for Mesher.set_surrogate: the largest |d_nn - d|
        over validation points (half uniform in the box, half within `near`
        of the surface) in each cell of a res^dim grid over the box and its
        neighbouring cells, times `safety`, at least `band` (e.g. 2 h); a
        cell without validation points gets the largest error overall.  For
        a network trained on a truncated SDF both values are clipped to
        +-trunc first, and a band of trunc or more becomes 1e30 (never
        skipped: the network's output cannot get past it).  Fine regions
        keep one bad spot of the network from closing a large block to it
        (on a 76-primitive CSG: 1 % of the interior decided without the
        exact program at res = 16, 27 % at 64, 38 % at 128 with 32 M
        validation points; each region's band also covers its neighbours'
        errors, so ~15-30 validation points per region sufficed there).
        Returns (bands (res,)*dim float32, lo, hi, largest error)."""
        if near is None:
            near = band
        x, d = self.sample(domain, self.bbox, nval, near, seed=seed, pad=0.0)
        dn = self.eval(x).astype(np.float64)
        if self.trunc:
            dn = np.clip(dn, -self.trunc, self.trunc)
            d = np.clip(d, -self.trunc, self.trunc)
        err = np.abs(dn - d)
        lo, hi = self.bbox[:, 0].copy(), self.bbox[:, 1].copy()
        idx = np.clip(((x - lo) / (hi - lo) * res).astype(np.int64), 0, res - 1)
        flat = np.ravel_multi_index(tuple(idx.T), (res,) * self.dim)
        cell = np.full(res ** self.dim, -1.0)
        np.maximum.at(cell, flat, err)
        emax = float(err.max())
        cell[cell < 0] = emax
        cell = cell.reshape((res,) * self.dim)
        # each cell takes its neighbours' errors too (points near a cell
        # face are judged by either cell)
        pad = np.pad(cell, 1, mode="edge")
        out = np.zeros_like(cell)
        for off in np.ndindex(*(3,) * self.dim):
            sl = tuple(slice(o, o + res) for o in off)
            out = np.maximum(out, pad[sl])
        bands = np.maximum(safety * out, band).astype(np.float32)
        if self.trunc:
            # a network trained on an SDF truncated at trunc never leaves
            # [-trunc, trunc]: where the band reaches trunc no decision is
            # safe there, and the band is made unreachable (the exact
            # program decides every evaluation in that region)
            bands[bands >= np.float32(self.trunc)] = np.float32(1e30)
        return bands, lo, hi, emax

    def max_error(self, domain, x):
        """max |d_nn - d| over points x, judged by the exact program (both
        clipped to +-trunc for a network trained on a truncated SDF)."""
        dn = self.eval(x).astype(np.float64)
        d = domain.eval(x)
        if self.trunc:
            dn = np.clip(dn, -self.trunc, self.trunc)
            d = np.clip(d, -self.trunc, self.trunc)
        return float(np.abs(dn - d).max())
