"""GPU kernels of the learned Schur guess against the CPU network reference
(oracle/mlp_ref.py), and the learned guess inside the IB solve.

Bars: the BF16 matrix-core GEMM equals the BF16-rounded product accumulated
in FP64 to relative 1e-6 (FP32 accumulation) for all four transpose
combinations (this also pins the MFMA operand layout); forward activations
agree to relative 1e-4 and parameter gradients to relative 2e-3 in L2 per
layer (activations that land on a BF16 rounding boundary may round
differently after a different FP32 summation order); training lowers the
loss; the learned guess converges to the zero-guess force within the
solver tolerance's reach (relative 1e-6) and the iteration counts of all
three starts are reported.
"""
import numpy as np
import pytest

from channel3d_parallel_amd import Channel, learned_guess
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    lay, total = mlp_ref.layout()
    P = np.zeros(total, dtype=np.float32)
    for (w, b, i, o) in lay:
        P[w:w + o * i] = rng.uniform(-1, 1, o * i) * np.sqrt(6.0 / i)
        P[b:b + o] = rng.uniform(-0.1, 0.1, o)
    return P


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_layout(ta, tb):
    print(f"\n[case] bf16 mfma gemm ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(4)
    M, N, K = 48, 40, 54
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    with Channel(18, 17, 18) as ch:
        C = ch.mlp_gemm(A, B, bool(ta), bool(tb))
    Am = mlp_ref.bf16(A.T if ta else A)
    Bm = mlp_ref.bf16(B.T if tb else B)
    ref = Am @ Bm
    assert _rel(C, ref) < 1e-6


def test_forward_and_gradients_vs_oracle():
    print("\n[case] mlp forward/backward vs oracle", flush=True)
    rng = np.random.default_rng(8)
    n = 96
    X = rng.standard_normal((n, 54)).astype(np.float32)
    T = rng.standard_normal((n, 4)).astype(np.float32)
    P = _params(2)
    loss_o, G_o = mlp_ref.loss_and_grad(P, X, T, bf16_ops=True)
    Y_o = mlp_ref.forward(P, X, True)[-1]
    with Channel(18, 17, 18) as ch:
        ch.mlp_init(1)
        ch.mlp_set_params(P)
        Y = ch.mlp_forward(X)
        G, loss = ch.mlp_gradients(X, T)
    assert _rel(Y, Y_o) < 1e-4
    assert abs(loss - loss_o) <= 1e-4 * loss_o
    lay, _ = mlp_ref.layout()
    for k, (w, b, i, o) in enumerate(lay):
        ew, eb = _rel(G[w:w + o * i], G_o[w:w + o * i]), _rel(G[b:b + o], G_o[b:b + o])
        print(f"  layer {k}: rel grad error W {ew:.2e} b {eb:.2e}")
        assert ew < 2e-3 and eb < 2e-3, k


def test_training_lowers_loss():
    print("\n[case] mlp training on a smooth synthetic map", flush=True)
    rng = np.random.default_rng(12)
    n = 4096
    X = rng.standard_normal((n, 54)).astype(np.float32)
    Wt = rng.standard_normal((4, 54)) / np.sqrt(54)
    T = np.tanh(X @ Wt.T).astype(np.float32)
    with Channel(18, 17, 18) as ch:
        ch.mlp_init(3)
        hist = ch.mlp_train(X, T, epochs=20, batch=256, lr=1e-3, seed=5)
        Y = ch.mlp_forward(X)
    print("  loss per epoch:", np.array2string(hist[[0, 4, 9, 19]], precision=3))
    assert hist[-1] < 0.1 * hist[0]
    assert np.mean((Y - T) ** 2) < 0.1 * np.mean(T ** 2)


def test_learned_guess_in_ib_solve():
    """Small two-cylinder case: train on pairs from two seeded runs, then
    compare zero, previous-force and learned starts on a held-out run."""
    print("\n[case] learned guess A/B, 34x33x34, 24x16 markers per cylinder", flush=True)
    nx, ny, nz, nth, nzb, dt = 34, 33, 34, 24, 16, 0.004
    X, T = learned_guess.generate_pairs(nx, ny, nz, nth, nzb, runs=2, steps=3, dt=dt, seed=11)
    s1, s2 = learned_guess.speeds(99, 0)
    ch, m = learned_guess.make_channel(nx, ny, nz, nth, nzb, s1, s2)
    with ch:
        hist, _, _ = learned_guess.train(ch, X, T, epochs=30, batch=256, lr=1e-3)
        rec = learned_guess.ab_compare(ch, dt, steps=2)
    print(f"  pairs {X.shape[0]}; loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    print(f"  iterations zero {rec['zero']} previous {rec['previous']} learned {rec['learned']}")
    print(f"  max |force diff| / max |force|: learned {max(rec['max_force_diff_learned_vs_zero']):.2e} "
          f"previous {max(rec['max_force_diff_previous_vs_zero']):.2e}")
    assert hist[-1] < hist[0]
    assert max(rec["max_force_diff_learned_vs_zero"]) < 1e-6
    assert sum(rec["learned"]) <= sum(rec["zero"])


def test_failed_buffer_growth_keeps_the_network_usable():
    """A forward pass whose activation buffers cannot be allocated (300 M rows,
    beyond HBM) returns C3D_ENOMEM and leaves the smaller buffer set that was
    in use intact: the next passes at the old and a larger size still run and
    give the same outputs."""
    from channel3d_parallel_amd._lib import fptr
    rng = np.random.default_rng(11)
    X = rng.standard_normal((1024, 54)).astype(np.float32)
    X2 = np.concatenate([X, rng.standard_normal((1024, 54)).astype(np.float32)])
    print("\n[case] network buffers: 1024 rows, 300M rows (fails), 1024, 2048", flush=True)
    with Channel(18, 17, 18) as ch:
        ch.mlp_init(seed=3)
        Y1 = ch.mlp_forward(X)
        Yb = np.zeros((1, 4), dtype=np.float32)
        rc = ch.L.c3d_mlp_forward(ch.h, fptr(X), 300_000_000, fptr(Yb))
        assert rc == -3, rc
        Y2 = ch.mlp_forward(X)
        Y3 = ch.mlp_forward(X2)
    assert np.array_equal(Y1, Y2)
    assert np.allclose(Y3[:1024], Y1, rtol=1e-6, atol=1e-6)
