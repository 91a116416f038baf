"""Oracle of the neural SDF surrogate (FP64 numpy).

Test infrastructure only (tests/ import it; the product never does).

The reference has no learned part (its SDFs are evaluated exactly,
SeismicMesh/geometry/signed_distance_functions.py:283-623), so this module
restates the surrogate's own definition:

* network: x' = (x - shift) * scale, a_0 = silu(W_in x' + b_in),
  a_l = silu(W_l a_{l-1} + b_l) (l = 1..L), y = w_out . a_L + b_out,
  d = out_scale * y;  silu(z) = z / (1 + exp(-z));
* loss over n samples with targets t (in the network's frame):
  mean (y_i - t_i)^2 + lam * mean (|g_i| - 1)^2 with the central
  difference g_i,k = (y(x'_i + h e_k) - y(x'_i - h e_k)) / (2 h), the
  eikonal term dropped for targets at a truncation |t_i| >= trunc;
* its gradient with respect to the flat parameter vector (layout of
  seismicmesh_amd.neural.param_layout), by hand-written backpropagation;
* the BF16 inference arithmetic: hidden weights and the activations that
  enter a matrix product rounded to BF16 (round to nearest even), products
  exact, sums in FP64 (the device accumulates in FP32).

Parity pins: the hand-written gradient is checked against central
differences of loss() in FP64 and against torch.autograd on the CPU
(tests/test_oracle_mlp.py); bf16_round against torch's bfloat16 cast.
"""
import numpy as np


def layout(dim, width, hidden):
    sizes = [("w_in", width * dim), ("b_in", width), ("w_hid", hidden * width * width),
             ("b_hid", hidden * width), ("w_out", width), ("b_out", 1)]
    out, off = {}, 0
    for name, n in sizes:
        out[name] = (off, n)
        off += n
    return out, off


def unpack(params, dim, width, hidden):
    lay, _ = layout(dim, width, hidden)
    g = {k: np.asarray(params[o:o + n], dtype=np.float64) for k, (o, n) in lay.items()}
    return dict(w_in=g["w_in"].reshape(width, dim), b_in=g["b_in"],
                w_hid=g["w_hid"].reshape(hidden, width, width),
                b_hid=g["b_hid"].reshape(hidden, width), w_out=g["w_out"], b_out=g["b_out"][0])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def make_inputs(x, shift, scale, eik_h):
    """Rows b * n + i: x'_i (b = 0), x'_i +- h e_k (b = 2k + 1, 2k + 2)."""
    x = np.asarray(x, dtype=np.float64)
    n, dim = x.shape
    q = (x - np.asarray(shift[:dim])) * np.asarray(scale[:dim])
    if eik_h == 0:
        return q
    blocks = [q]
    for k in range(dim):
        for sgn in (1.0, -1.0):
            e = q.copy()
            e[:, k] += sgn * eik_h
            blocks.append(e)
    return np.vstack(blocks)


def forward(P, X):
    Z, A = [], []
    z = X @ P["w_in"].T + P["b_in"]
    Z.append(z)
    A.append(silu(z))
    for l in range(P["w_hid"].shape[0]):
        z = A[-1] @ P["w_hid"][l].T + P["b_hid"][l]
        Z.append(z)
        A.append(silu(z))
    y = A[-1] @ P["w_out"] + P["b_out"]
    return y, Z, A


def loss_terms(y, t, n, dim, lam, eik_h, trunc=0.0):
    e = y[:n] - t
    L = float(np.mean(e * e))
    gy = np.zeros_like(y)
    gy[:n] = 2.0 * e / n
    if lam > 0:
        g = np.stack([(y[(2 * k + 1) * n:(2 * k + 2) * n] - y[(2 * k + 2) * n:(2 * k + 3) * n])
                      / (2.0 * eik_h) for k in range(dim)], axis=1)
        gn = np.sqrt((g * g).sum(axis=1))
        r = gn - 1.0
        # samples at the truncation have no eikonal term
        on = np.ones(n) if trunc <= 0 else (np.abs(t) < trunc).astype(float)
        L += lam * float(np.sum(on * r * r) / n)
        c = on * np.where(gn > 0, lam * 2.0 * r / np.where(gn > 0, gn, 1.0) / n / (2.0 * eik_h),
                          0.0)
        for k in range(dim):
            gy[(2 * k + 1) * n:(2 * k + 2) * n] = c * g[:, k]
            gy[(2 * k + 2) * n:(2 * k + 3) * n] = -c * g[:, k]
    return L, gy


def loss(params, dim, width, hidden, x, t, shift, scale, lam=0.1, eik_h=5e-3, trunc=0.0):
    P = unpack(params, dim, width, hidden)
    X = make_inputs(x, shift, scale, eik_h if lam > 0 else 0.0)
    y, _, _ = forward(P, X)
    return loss_terms(y, np.asarray(t, dtype=np.float64), len(x), dim, lam, eik_h, trunc)[0]


def loss_and_grad(params, dim, width, hidden, x, t, shift, scale, lam=0.1, eik_h=5e-3,
                  trunc=0.0):
    """(loss, gradient) with the gradient in the flat parameter layout."""
    P = unpack(params, dim, width, hidden)
    X = make_inputs(x, shift, scale, eik_h if lam > 0 else 0.0)
    y, Z, A = forward(P, X)
    L, gy = loss_terms(y, np.asarray(t, dtype=np.float64), len(x), dim, lam, eik_h, trunc)
    grads = {"w_out": A[-1].T @ gy, "b_out": np.array([gy.sum()])}
    dA = np.outer(gy, P["w_out"])
    gw = np.zeros_like(P["w_hid"])
    gb = np.zeros_like(P["b_hid"])
    for l in range(hidden, 0, -1):
        dZ = dA * dsilu(Z[l])
        gw[l - 1] = dZ.T @ A[l - 1]
        gb[l - 1] = dZ.sum(axis=0)
        dA = dZ @ P["w_hid"][l - 1]
    dZ = dA * dsilu(Z[0])
    grads["w_in"] = dZ.T @ X
    grads["b_in"] = dZ.sum(axis=0)
    grads["w_hid"] = gw
    grads["b_hid"] = gb
    lay, total = layout(dim, width, hidden)
    flat = np.zeros(total)
    for k, (o, n) in lay.items():
        flat[o:o + n] = np.asarray(grads[k]).ravel()
    return L, flat


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned
    as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    out = ((u + r) & 0xFFFF0000).astype(np.uint32)
    return out.view(np.float32)


def forward_bf16(params, dim, width, hidden, x, shift, scale, out_scale):
    """The inference kernel's arithmetic (see module docstring)."""
    P = unpack(params, dim, width, hidden)
    q = ((np.asarray(x, dtype=np.float64) - np.asarray(shift[:dim])) *
         np.asarray(scale[:dim])).astype(np.float32).astype(np.float64)
    a = bf16_round(silu(q @ P["w_in"].T + P["b_in"])).astype(np.float64)
    for l in range(hidden):
        w = bf16_round(P["w_hid"][l]).astype(np.float64)
        a = bf16_round(silu(a @ w.T + P["b_hid"][l])).astype(np.float64)
    y = a @ P["w_out"] + P["b_out"]
    return out_scale * y
