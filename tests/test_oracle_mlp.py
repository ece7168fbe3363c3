"""Pins of the CPU network reference (oracle/mlp_ref.py): FP64 backward pass
against central finite differences, BF16 rounding against its definition."""
import numpy as np

from oracle import mlp_ref


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    lay, total = mlp_ref.layout()
    P = np.zeros(total)
    for (w, b, i, o) in lay:
        P[w:w + o * i] = rng.uniform(-1, 1, o * i) * np.sqrt(6.0 / i) * scale
        P[b:b + o] = rng.uniform(-0.1, 0.1, o)
    return P


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = _params(1)
    X = rng.standard_normal((24, 54))
    T = rng.standard_normal((24, 4))
    _, G = mlp_ref.loss_and_grad(P, X, T, bf16_ops=False)
    lay, total = mlp_ref.layout()
    idx = np.concatenate([rng.integers(w, w + o * i, 6) for (w, b, i, o) in lay] +
                         [rng.integers(b, b + o, 2) for (w, b, i, o) in lay])
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (mlp_ref.loss_and_grad(Pp, X, T, False)[0] - mlp_ref.loss_and_grad(Pm, X, T, False)[0]) / (2 * h)
        assert abs(fd - G[k]) <= 1e-6 * max(1.0, abs(G[k])), (k, fd, G[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-5, 0.0], dtype=np.float32)
    r = mlp_ref.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 3*2^-9 rounds up
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5 and r[5] == 0.0
    # relative error bound: half a BF16 ulp = 2^-8 (8 significant bits)
    assert abs(r[4] - 3.0e-5) <= 2 ** -8 * 3.0e-5
    y = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    assert np.all(np.abs(mlp_ref.bf16(y) - y) <= 2 ** -8 * np.abs(y) + 1e-45)
