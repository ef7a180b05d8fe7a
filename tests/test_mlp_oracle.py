"""The CPU restatement of the warm-start MLP (oracle/mlp_oracle.py), pinned
by mathematics: its analytic gradient against central finite differences in
FP64, its BF16 rounding against closed-form cases."""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_closed_forms():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing just above 1
    vals = np.array([one, one + ulp / 2, one + 1.5 * ulp, one + ulp / 2 + 2 ** -20, -one - 1.5 * ulp,
                     (one + ulp / 2 + ulp / 4) * 2.0 ** 100, 2.0 ** -130])
    got = mo.bf16_round(vals)
    exp = np.array([one, one, one + 2 * ulp, one + ulp, -one - 2 * ulp, (one + ulp) * 2.0 ** 100, 2.0 ** -130])
    assert got[0] == exp[0] and got[1] == exp[1], "ties go to even"
    assert got[2] == exp[2] and got[4] == exp[4], "ties go to even (odd neighbour)"
    assert got[3] == exp[3], "above the tie rounds up"
    assert got[5] == exp[5], "large magnitude rounds up past the tie"
    assert got[6] == exp[6], "exactly representable subnormal"


def _random_net(seed, dims):
    rng = np.random.default_rng(seed)
    n = sum(fi * fo + fo for fi, fo in mo.layer_shapes(*dims))
    p = rng.normal(0, 0.4, n)
    x = rng.uniform(-1, 1, (37, dims[0]))
    y = rng.normal(0, 1, (37, dims[3]))
    return p, x, y


def test_exact_gradient_matches_finite_differences():
    dims = (4, 16, 3, 4)
    p, x, y = _random_net(1, dims)
    loss, g = mo.grad_exact(p, x, y, dims)
    rng = np.random.default_rng(2)
    idx = rng.choice(len(p), 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.grad_exact(p + e, x, y, dims)[0] - mo.grad_exact(p - e, x, y, dims)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_emulated_close_to_exact_and_shapes():
    dims = (4, 64, 5, 4)
    p, x, y = _random_net(3, dims)
    l0, g0 = mo.grad_exact(p, x, y, dims)
    l1, g1 = mo.loss_grad(p, x, y, dims)
    assert g0.shape == g1.shape == (4 * 64 + 64 + 4 * (64 * 64 + 64) + 64 * 4 + 4,)
    assert abs(l1 - l0) / l0 < 0.05
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.1
