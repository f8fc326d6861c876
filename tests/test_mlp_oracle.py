"""Pins of the wall-network oracle (oracle/mlp.py) on the CPU: BF16 rounding
against torch's conversion, SiLU and its derivative, the backward pass against
central finite differences, Adam's first step in closed form, and the
wall-cell features/targets against the log law."""
import numpy as np
import pytest

from oracle import mlp as om


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 20000),
                        np.array([0.0, -0.0, 1.0, -1.0, 1.00390625, 1.01171875, 3.0e38], np.float32)])
    x = x.astype(np.float32)
    # ties: 1 + 2^-8 lies half-way between 1 and 1 + 2^-7
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(om.bf16(x), ref)
    assert om.bf16(np.float32(1.00390625)) == 1.0


def test_silu_and_derivative():
    z = np.linspace(-12, 12, 2001)
    assert om.silu(0.0) == 0.0 and om.dsilu(0.0) == 0.5
    h = 1e-6
    fd = (om.silu(z + h) - om.silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - om.dsilu(z))) < 1e-8
    # large positive z: silu -> z, derivative -> 1
    assert abs(om.silu(40.0) - 40.0) < 1e-12 and abs(om.dsilu(40.0) - 1.0) < 1e-12


def _random_net(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(om.NPARAM)
    for off, fan, cnt in ((om.W1, 4, 256), (om.W2, 64, 4096), (om.W3, 64, 4096), (om.W4, 64, 128)):
        p[off:off + cnt] = rng.uniform(-1, 1, cnt) * np.sqrt(6.0 / fan)
    p[om.B1:om.W2] = rng.normal(0, 0.1, 64)
    p[om.B2:om.W3] = rng.normal(0, 0.1, 64)
    p[om.B3:om.W4] = rng.normal(0, 0.1, 64)
    p[om.B4:] = rng.normal(0, 0.1, 2)
    X = rng.normal(0, 1, (37, 4))
    T = rng.normal(0, 1, (37, 2))
    norm = np.array([0.1, -0.2, 0.3, 0.0, 1.5, 0.7, 1.0, 2.0, 0.2, -0.1, 1.3, 0.8])
    return p, norm, X, T


def test_backward_against_finite_differences():
    p, norm, X, T = _random_net(1)
    g, loss = om.gradients(p, norm, X, T, quant=False)
    assert abs(loss - om.loss_only(p, norm, X, T)) < 1e-14
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.integers(lo, hi, 6) for lo, hi in
                          ((om.W1, om.B1), (om.B1, om.W2), (om.W2, om.B2), (om.B2, om.W3),
                           (om.W3, om.B3), (om.B3, om.W4), (om.W4, om.B4), (om.B4, om.NPARAM))])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (om.loss_only(pp, norm, X, T) - om.loss_only(pm, norm, X, T)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])), (q, fd, g[q])


def test_quantised_gradient_is_close_to_exact():
    # the BF16 rounding points perturb the gradient by O(2^-8), not more
    p, norm, X, T = _random_net(3)
    g0, _ = om.gradients(p, norm, X, T, quant=False)
    g1, _ = om.gradients(p, norm, X, T, quant=True)
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.05


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m1, v1 = om.adam_step(p, np.zeros(100), np.zeros(100), g, 1, lr=1e-3)
    # t = 1: m_hat = g, v_hat = g^2, step = lr g / (|g| + eps)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    # t = 2 with the same gradient: m_hat = v_hat^(1/2) = |g| again
    p2, _, _ = om.adam_step(p1, m1, v1, g, 2, lr=1e-3)
    assert np.allclose(p2, p1 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)


def test_wall_samples_follow_the_log_law():
    import oracle
    nx, ny, nz = 10, 9, 10
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    y = np.concatenate([[0.0], np.linspace(0.0, 2.0, ny)])
    yg = np.zeros(nyg + 1)
    yg[2:nyg] = 0.5 * (y[1:ny] + y[2:ny + 1])
    yg[1] = 2 * y[1] - yg[2]
    yg[nyg] = 2 * y[ny] - yg[nyg - 1]
    U = np.full((nx, nyg, nzg), 0.8)
    W = np.zeros((nxg, nyg, nz))
    alpha = np.zeros((nx, 2, nzg))
    nu, kappa, B = 1e-4, 0.41, 5.2
    X, Y = om.wall_samples(U, W, alpha, y, yg, nu, 0.3, kappa, B, 0.0, oracle.loglaw_utau)
    h = 0.3
    ut = oracle.loglaw_utau(0.8, h, nu, kappa, B)
    yp = h * ut / nu
    assert abs(0.8 / ut - (np.log(yp) / kappa + B)) < 1e-10
    assert np.allclose(X[:, 0], np.arcsinh(h * 0.8 / nu / 10))
    assert np.allclose(X[:, 1:], 0.0)
    assert np.allclose(Y[:, 0], np.log1p(yp * yp)) and np.all(Y[:, 1] == 0.0)
