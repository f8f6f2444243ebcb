"""Pins of the FP64 MLP oracle: its gradients are the derivative of its
loss (central finite differences), and its bf16 rounding is round-to-nearest
-even."""
import numpy as np


def test_mlp_gradient_is_the_loss_derivative():
    """Central finite differences of the exact FP64 loss pin the oracle's
    backward formulas (the same code runs with the bf16 roundings for the
    GPU comparison)."""
    from oracle import mlp as M
    rng = np.random.default_rng(3)
    shape = (12, 16, 16, 10)
    W = [M.bf16(rng.normal(0, 0.5, size=(shape[i], shape[i + 1]))).astype(np.float32)
         for i in range(3)]
    b = [rng.normal(0, 0.1, size=shape[i + 1]).astype(np.float32) for i in range(3)]
    X = M.bf16(rng.normal(size=(9, shape[0]))).astype(np.float32)
    y = rng.integers(0, shape[3], size=9)
    gW, gb, _ = M.grads(W, b, X, y, q=M._exact)
    for l in range(3):
        for _ in range(4):
            i = rng.integers(0, shape[l])
            j = rng.integers(0, shape[l + 1])
            # a step of 2^-4 on a bf16-representable weight stays on the grid
            h = 1e-5
            Wp = [w.astype(np.float64) for w in W]
            Wm = [w.astype(np.float64) for w in W]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fp = M.grads(Wp, b, X, y, q=M._exact)[2]
            fm = M.grads(Wm, b, X, y, q=M._exact)[2]
            num = (fp - fm) / (2 * h)
            assert abs(num - gW[l][i, j]) < 1e-6 + 1e-5 * abs(num), (l, i, j, num, gW[l][i, j])
    # bias gradients the same way
    for l in range(3):
        j = rng.integers(0, shape[l + 1])
        bp = [x.astype(np.float64).copy() for x in b]
        bm = [x.astype(np.float64).copy() for x in b]
        bp[l][j] += 1e-5
        bm[l][j] -= 1e-5
        num = (M.grads(W, bp, X, y, q=M._exact)[2] - M.grads(W, bm, X, y, q=M._exact)[2]) / 2e-5
        assert abs(num - gb[l][j]) < 1e-6 + 1e-5 * abs(num)



def test_bf16_rounding_nearest_even():
    from oracle import mlp as M
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 3.140625], np.float32)
    # 1+2^-8 is halfway between 1 and 1+2^-7: ties to even (1.0); 1+3*2^-8
    # is halfway between 1+2^-7 and 1+2^-6: ties to even (1+2^-6)
    assert list(M.bf16(x)) == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -2.5, 3.140625]
