"""CPU: the numpy MLP oracle's backward is the derivative of its forward
(central differences in FP64 without bf16 rounding), and bf16 rounding is
round-to-nearest-even."""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], np.float32)
    r = mo.bf16(x)
    np.testing.assert_array_equal(r[:3], [1.0, 1.0, 1.0 + 2 ** -6])  # ties to even
    assert r[3] == -2.5
    assert abs(r[4] / 1e-3 - 1) < 2 ** -8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    W, b = mo.init_params(3)
    W = [w.astype(np.float64) for w in W]
    b = [(0.1 * rng.standard_normal(x.shape)).astype(np.float64) for x in b]
    X = rng.standard_normal((24, 32))
    t = rng.standard_normal(24)

    def loss(Wl, bl):
        y, _, _ = mo.forward(X, Wl, bl, round_bf16=False)
        return 0.5 * np.mean((y - t) ** 2)

    L0, gW, gb = mo.backward(X, t, W, b, round_bf16=False)
    # forward in float32 -> use a coarse step and relative tolerance
    for li, (i, j) in [(0, (3, 5)), (1, (7, 100)), (2, (127, 0)), (3, (17, 0))]:
        h = 1e-2
        Wp = [w.copy() for w in W]; Wp[li][i, j] += h
        Wm = [w.copy() for w in W]; Wm[li][i, j] -= h
        fd = (loss(Wp, b) - loss(Wm, b)) / (2 * h)
        assert abs(fd - gW[li][i, j]) <= 2e-3 * max(abs(fd), 1e-3), (li, fd, gW[li][i, j])
    for li, j in [(0, 9), (2, 60), (3, 0)]:
        h = 1e-2
        bp = [x.copy() for x in b]; bp[li][j] += h
        bm = [x.copy() for x in b]; bm[li][j] -= h
        fd = (loss(W, bp) - loss(W, bm)) / (2 * h)
        assert abs(fd - gb[li][j]) <= 2e-3 * max(abs(fd), 1e-3)


def test_pack_roundtrip():
    W, b = mo.init_params(5)
    p = mo.pack(W, b)
    assert p.size == 32 * 128 + 128 + 2 * (128 * 128 + 128) + 128 + 1
    W2, b2 = mo.unpack(p)
    for x, y in zip(W + b, W2 + b2):
        np.testing.assert_array_equal(x, y)
