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
@pytest.mark.parametrize("M,N,K", [(48, 40, 54), (1100, 1030, 70), (256, 256, 256), (1024, 1024, 128)])
def test_gemm_layout(ta, tb, M, N, K):
    """Both tile sizes: 32 x 32 for the small products, 64 x 64 when the
    64-tile grid has at least 256 workgroups (18 x 17 and 16 x 16 here);
    ragged edges in M, N and K; scalar operand loads where an extent is not a
    multiple of 4, float4 loads where it is (all of the last two shapes, some
    operands of the first two)."""
    print(f"\n[case] bf16 mfma gemm ta={ta} tb={tb} {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(4)
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


def test_learned_guess_two_cylinders_128():
    """The learned-start workload itself: two counter-rotating cylinders in
    the 128x128x128-cell channel (130x129x130 faces, uniform y), 96 x 128
    markers per cylinder (nb = 24,576).  20k training pairs from two runs
    with rotation speeds drawn from U[0.3, 0.7] U_b, the network trained on
    the GPU by the library's kernels, then, on a held-out run, every start
    solved from the same states to the same tolerance.  The learned start's
    forces equal the zero-guess forces to the solver's accuracy; iterations
    per RK3 stage of every start are printed (the counts are measurements,
    not asserted: no start is faster than zero by a fixed factor)."""
    nx, ny, nz, nth, nzb, dt = 130, 129, 130, 96, 128, 0.01
    print(f"\n[case] learned guess, {nx}x{ny}x{nz} faces, {nth}x{nzb} markers per cylinder", flush=True)
    X, T = learned_guess.generate_pairs(nx, ny, nz, nth, nzb, runs=2, steps=3, dt=dt, seed=11,
                                        max_pairs=20000, warm_steps=2)
    assert X.shape == (20000, 54) and T.shape == (20000, 4)
    s1, s2 = learned_guess.speeds(99, 0)
    ch, m = learned_guess.make_channel(nx, ny, nz, nth, nzb, s1, s2)
    assert m.nb == 2 * 96 * 128
    with ch:
        for _ in range(2):
            ch.compute_time_step_RK3(dt)
        hist, isc, osc = learned_guess.train(ch, X, T, epochs=60, batch=256, lr=1e-3)
        sch, lch = learned_guess.side_channels(ch, nx, ny, nz, nth, nzb, s1, s2, isc, osc)
        with sch, lch:
            rec = learned_guess.ab_compare(ch, dt, steps=2, learned=lch, stage_ch=sch)
    print(f"  loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    for n in ("zero", "previous", "stage", "learned", "history"):
        print(f"  {n:9s} per stage {learned_guess.per_stage(rec, n)} total {sum(rec[n])}")
    diffs = {n: max(rec[f"max_force_diff_{n}_vs_zero"]) for n in ("previous", "stage", "learned", "history")}
    print(f"  max |force - zero-guess force| / max |force|: {diffs}")
    assert hist[-1] < 0.5 * hist[0]
    # every start converges to the zero-guess force (absolute residual 1e-10)
    for n, d in diffs.items():
        assert d < 1e-8, (n, d)


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
