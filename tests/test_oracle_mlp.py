"""Pin of the FP64 MLP reference: backward pass against central finite
differences, softplus/sigmoid identities."""
import numpy as np


def test_mlp_gradients_match_finite_differences():
    from oracle import mlp_oracle as M
    sizes = (7, 12, 10, 9, 6)
    p = M.init_params(1, sizes)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(5, 7))
    t = rng.normal(size=(5, 6))
    _, g = M.loss_and_grads(p, x, t)
    h = 1e-6
    for i in range(len(p)):
        for j in range(2):
            arr = p[i][j]
            it = np.nditer(arr, flags=["multi_index"])
            for cnt, _ in enumerate(it):
                if cnt > 12:
                    break
                idx = it.multi_index
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = M.loss_and_grads(p, x, t)
                arr[idx] = old - h
                lm, _ = M.loss_and_grads(p, x, t)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[i][j][idx]) < 1e-7 + 1e-5 * abs(fd)


def test_softplus_derivative_is_sigmoid():
    from oracle import mlp_oracle as M
    z = np.linspace(-30, 30, 121)
    h = 1e-6
    fd = (M.softplus(z + h) - M.softplus(z - h)) / (2 * h)
    assert np.allclose(fd, M.sigmoid(z), atol=1e-8)
    assert np.allclose(M.softplus(z) - M.softplus(-z), z, atol=1e-12)
