"""TEST INFRASTRUCTURE ONLY -- CPU restatement (numpy, FP64) of the warm-start
MLP of the package (include/psst_mlp.h): used by the tests as the checker of
the GPU forward and backward kernels, never by the product path.

The reference has no learned component (it starts every solve from the
.par freestream or a .sav restart, psstFiniteVolumeGrid.cpp:641-702), so
this restatement is pinned by mathematics instead of reference vectors:
its analytic gradient is checked against central finite differences in
FP64 (`grad_exact`), and its BF16 rounding against closed-form cases.

Two forms:
  * exact:    FP64 throughout (the mathematics of the network);
  * emulated: the GPU's rounding points -- BF16 weights, BF16 hidden
    activations H and hidden-layer gradients dZ (round to nearest even from
    FP32) -- with FP64 accumulation; the GPU differs from it only by FP32
    accumulation order.
Parameter layout: for each layer, W [fan_in][fan_out] then b [fan_out].
"""
import numpy as np


def bf16_round(a):
    """Round to BF16 (nearest, ties to even) via FP32, as float64 values."""
    f = np.asarray(a, dtype=np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    bias = ((bits >> 16) & 1) + 0x7FFF
    r = ((bits + bias) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def layer_shapes(n_in, width, depth, n_out):
    dims = [n_in] + [width] * depth + [n_out]
    return [(dims[l], dims[l + 1]) for l in range(depth + 1)]


def unpack(params, n_in, width, depth, n_out):
    out, off = [], 0
    for fi, fo in layer_shapes(n_in, width, depth, n_out):
        W = np.asarray(params[off:off + fi * fo], dtype=np.float64).reshape(fi, fo)
        off += fi * fo
        b = np.asarray(params[off:off + fo], dtype=np.float64)
        off += fo
        out.append((W, b))
    assert off == len(params)
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _run(params, x, y, dims, emulate):
    n_in, width, depth, n_out = dims
    rnd = bf16_round if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    layers = [(rnd(W), rnd(b)) for W, b in unpack(params, *dims)]
    x = np.asarray(x, dtype=np.float64)
    Zs, Hs = [], []
    h = x
    for l in range(depth):
        W, b = layers[l]
        z = h @ W + b
        Zs.append(z)
        h = rnd(silu(z))
        Hs.append(h)
    W, b = layers[depth]
    out = h @ W + b
    if y is None:
        return out, None, None
    y = np.asarray(y, dtype=np.float64)
    d = out - y
    loss = float(np.mean(d * d))
    dout = 2.0 * d / d.size
    grads = [None] * (depth + 1)
    grads[depth] = (Hs[-1].T @ dout, dout.sum(axis=0))
    dz = rnd((dout @ W.T) * dsilu(Zs[-1]))
    for l in range(depth - 1, 0, -1):
        W, _ = layers[l]
        grads[l] = (Hs[l - 1].T @ dz, dz.sum(axis=0))
        dz = rnd((dz @ W.T) * dsilu(Zs[l - 1]))
    grads[0] = (x.T @ dz, dz.sum(axis=0))
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return out, loss, flat


def forward(params, x, dims, emulate=True):
    return _run(params, x, None, dims, emulate)[0]


def loss_grad(params, x, y, dims, emulate=True):
    _, loss, g = _run(params, x, y, dims, emulate)
    return loss, g


def grad_exact(params, x, y, dims):
    return loss_grad(params, x, y, dims, emulate=False)
