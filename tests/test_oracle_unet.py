"""CPU checks of the U-Net oracle (oracle/unet.py): its analytic gradient against central finite
differences in FP64, and closed forms of the residual-norm loss.  These pin the oracle the GPU
network is compared with."""
import numpy as np
import pytest

import unet


def fd_error(p, d, rhs, solid, flags, g, mode="residual"):
    """Smallest relative error of central differences at two step sizes against g . d (a ReLU input
    that changes sign inside the larger step spoils that one; a wrong gradient fails both)."""
    errs = []
    for h in (1e-7, 1e-8):
        lp = unet.loss(*unet.forward(p + h * d, rhs, solid)[:2], flags, mode)[0]
        lm = unet.loss(*unet.forward(p - h * d, rhs, solid)[:2], flags, mode)[0]
        fd = (lp - lm) / (2 * h)
        errs.append(abs(fd - g @ d) / max(abs(fd), 1e-12))
    return min(errs)


def scene(n, seed, obstacles=True):
    rng = np.random.default_rng(seed)
    rhs = rng.normal(size=(n, n, n))
    solid = None
    if obstacles:
        solid = np.zeros((n, n, n), dtype=np.uint8)
        solid[1:3, 2:4, 5:7] = 1
        rhs[solid == 1] = 0.0
    return rhs, solid


@pytest.mark.parametrize("mode", ["residual", "energy"])
@pytest.mark.parametrize("obstacles", [False, True])
def test_gradient_matches_finite_differences(obstacles, mode):
    n = 8
    rhs, solid = scene(n, 3, obstacles)
    p = unet.init_params(seed=5).astype(np.float64)
    flags = unet.fluid_flags(solid, rhs.shape)
    y, cache = unet.forward(p, rhs, solid)
    L0, dy = unet.loss(y, cache, flags, mode)
    g = unet.backward(p, cache, dy)
    rng = np.random.default_rng(11)
    for _ in range(3):
        d = rng.normal(size=p.size)
        assert fd_error(p, d, rhs, solid, flags, g, mode) < 1e-6


def test_gradient_of_single_blocks():
    # restricted to one parameter block at a time: the output layer, a decoder conv with a concat
    # input, an encoder conv behind a pooling, and the first conv
    n = 8
    rhs, solid = scene(n, 4)
    p = unet.init_params(seed=6).astype(np.float64)
    flags = unet.fluid_flags(solid, rhs.shape)
    y, cache = unet.forward(p, rhs, solid)
    _, dy = unet.loss(y, cache, flags)
    g = unet.backward(p, cache, dy)
    offs, oo, nparam = unet.offsets()
    rng = np.random.default_rng(2)
    for lo, hi in [(oo, nparam), offs[8], offs[12], offs[4], (offs[0][0], offs[0][1])]:
        d = np.zeros(nparam)
        d[lo:hi] = rng.normal(size=hi - lo)
        assert fd_error(p, d, rhs, solid, flags, g) < 1e-6


def test_loss_is_zero_at_the_solution():
    # build b = L x for a known x: the network's loss of y = x is 0, and of y = x + const too
    # (constants are the Neumann null space)
    n = 8
    rng = np.random.default_rng(1)
    x = rng.normal(size=(n, n, n))
    flags = unet.fluid_flags(None, x.shape)
    b = unet.laplacian(x, flags)
    cache = {"b": b, "bf": False}
    assert unet.loss(x, cache, flags)[0] < 1e-24
    assert unet.loss(x + 3.0, cache, flags)[0] < 1e-20
    # a wrong answer has the expected loss: y = 0 gives exactly 1
    assert abs(unet.loss(np.zeros_like(x), cache, flags)[0] - 1.0) < 1e-14


def test_energy_loss_is_the_a_norm_error():
    # energy(y) - energy(x) = |y - x|_A^2 / (2 |b|^2) for the solution x of L x = b
    n = 8
    rng = np.random.default_rng(2)
    x = rng.normal(size=(n, n, n))
    flags = unet.fluid_flags(None, x.shape)
    b = unet.laplacian(x, flags)
    cache = {"b": b, "bf": False}
    y = x + 0.3 * rng.normal(size=x.shape)
    e = y - x
    lhs = unet.loss(y, cache, flags, "energy")[0] - unet.loss(x, cache, flags, "energy")[0]
    assert abs(lhs - (e * unet.laplacian(e, flags)).sum() / (2 * (b ** 2).sum())) < 1e-12


def test_laplacian_matches_assembled_matrix():
    # the stencil with obstacle walls equals the explicitly assembled graph Laplacian
    n = 4
    solid = np.zeros((n, n, n), dtype=np.uint8)
    solid[1, 1, 1] = 1
    solid[2, 3, 0] = 1
    flags = unet.fluid_flags(solid, solid.shape)
    idx = np.arange(n ** 3).reshape(n, n, n)
    A = np.zeros((n ** 3, n ** 3))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                if solid[i, j, k]:
                    continue
                for di, dj, dk in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
                    a, b, c = i + di, j + dj, k + dk
                    if 0 <= a < n and 0 <= b < n and 0 <= c < n and not solid[a, b, c]:
                        A[idx[i, j, k], idx[i, j, k]] += 1
                        A[idx[i, j, k], idx[a, b, c]] -= 1
    y = np.random.default_rng(0).normal(size=(n, n, n))
    assert np.allclose(unet.laplacian(y, flags).reshape(-1), A @ y.reshape(-1), atol=1e-12)


def test_bf16_rounding():
    # nearest even on the FP32 bits: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> 1; 1 + 3 * 2^-8 -> 1 + 2^-6
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0])
    assert np.array_equal(unet.bf16(v), np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 0.0]))


def test_parameter_layout():
    offs, oo, n = unet.offsets()
    assert n == 684817
    assert offs[1][0] == 27 * 16 * 16 + 16
    assert oo == n - 17
