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


# ------------------------------------------------------------------------
# The forward pass, its gradients and the fallback stress pinned by an
# independent implementation (torch, float64) and by closed forms.

def _layout():
    """Offsets of W1 b1 W2 b2 W3 b3 W4 b4 from the documented sizes
    (W[k][n]: input k, output n): 4x64, 64, 64x64, 64, 64x64, 64, 64x2, 2."""
    sizes = [(4, 64), (64,), (64, 64), (64,), (64, 64), (64,), (64, 2), (2,)]
    offs, o = [], 0
    for s in sizes:
        offs.append((o, s))
        o += int(np.prod(s))
    assert o == 8770
    return offs


def _torch_model(p, torch):
    lay = _layout()
    m = torch.nn.Sequential(torch.nn.Linear(4, 64), torch.nn.SiLU(), torch.nn.Linear(64, 64), torch.nn.SiLU(),
                            torch.nn.Linear(64, 64), torch.nn.SiLU(), torch.nn.Linear(64, 2)).double()
    lin = [m[0], m[2], m[4], m[6]]
    with torch.no_grad():
        for q, layer in enumerate(lin):
            (ow, sw), (ob, sb) = lay[2 * q], lay[2 * q + 1]
            W = torch.tensor(p[ow:ow + int(np.prod(sw))].reshape(sw))
            layer.weight.copy_(W.T)       # Linear computes x @ weight.T
            layer.bias.copy_(torch.tensor(p[ob:ob + sb[0]]))
    return m


def test_forward_matches_torch_float64():
    torch = pytest.importorskip("torch")
    p, norm, X, _ = _random_net(5)
    m = _torch_model(p, torch)
    x = (torch.tensor(X.astype(np.float32).astype(np.float64)) - torch.tensor(norm[0:4])) / torch.tensor(norm[4:8])
    ref = (m(x) * torch.tensor(norm[10:12]) + torch.tensor(norm[8:10])).detach().numpy()
    got = om.forward(p, norm, X, quant=False)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_gradients_match_torch_autograd():
    torch = pytest.importorskip("torch")
    p, norm, X, T = _random_net(6)
    m = _torch_model(p, torch)
    x = (torch.tensor(X.astype(np.float32).astype(np.float64)) - torch.tensor(norm[0:4])) / torch.tensor(norm[4:8])
    t = (torch.tensor(T.astype(np.float32).astype(np.float64)) - torch.tensor(norm[8:10])) / torch.tensor(norm[10:12])
    loss = 0.5 * ((m(x) - t) ** 2).sum() / len(X)
    loss.backward()
    g, lo = om.gradients(p, norm, X, T, quant=False)
    assert abs(lo - loss.item()) <= 1e-13 * loss.item()
    lay = _layout()
    lin = [m[0], m[2], m[4], m[6]]
    for q, layer in enumerate(lin):
        (ow, sw), (ob, sb) = lay[2 * q], lay[2 * q + 1]
        gw = layer.weight.grad.T.numpy().ravel()
        gb = layer.bias.grad.numpy()
        assert np.max(np.abs(g[ow:ow + gw.size] - gw)) <= 1e-12 * np.max(np.abs(gw)), ("W", q + 1)
        assert np.max(np.abs(g[ob:ob + gb.size] - gb)) <= 1e-12 * np.max(np.abs(gb)), ("b", q + 1)


def test_quantised_forward_matches_torch_bfloat16_points():
    # the BF16 rounding points of the quantised forward (weights, h1..h3),
    # rounded by torch's own conversion
    torch = pytest.importorskip("torch")
    p, norm, X, _ = _random_net(7)
    lay = _layout()

    def b(a):
        return a.to(torch.bfloat16).to(torch.float64)

    def W(q):
        (o, s) = lay[2 * q]
        return torch.tensor(p[o:o + int(np.prod(s))].reshape(s))

    def bias(q):
        (o, s) = lay[2 * q + 1]
        return torch.tensor(p[o:o + s[0]])

    h = (torch.tensor(X.astype(np.float32).astype(np.float64)) - torch.tensor(norm[0:4])) / torch.tensor(norm[4:8])
    for q in range(3):
        h = b(torch.nn.functional.silu(h @ b(W(q)) + bias(q)))
    y = (h @ b(W(3)) + bias(3)) * torch.tensor(norm[10:12]) + torch.tensor(norm[8:10])
    got = om.forward(p, norm, X, quant=True)
    assert np.max(np.abs(got - y.numpy())) <= 1e-12 * np.max(np.abs(y.numpy()))


def test_quantised_forward_hand_computed():
    # W2 = W3 = 0, b3 = 1, W4 = 1/2: h3 = bf16(silu(1)) = 187/256 for every
    # input (silu(1) = 0.7310585786..., BF16 keeps 8 significant bits), so
    # y_n = 64 * 187/256 * 1/2 + b4_n exactly
    p, _, X, _ = _random_net(8)
    lay = _layout()
    for q in (1, 2):
        o, s = lay[2 * q]
        p[o:o + int(np.prod(s))] = 0.0
    o, s = lay[5]
    p[o:o + 64] = 1.0
    o, s = lay[6]
    p[o:o + 128] = 0.5
    o, s = lay[7]
    p[o:o + 2] = [0.25, -3.0]
    norm = np.array([0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1], float)
    got = om.forward(p, norm, X, quant=True)
    assert np.all(got[:, 0] == 64 * 187 / 256 * 0.5 + 0.25)
    assert np.all(got[:, 1] == 64 * 187 / 256 * 0.5 - 3.0)


def test_fallback_stress_closed_form():
    # equilibrium_alpha_2d: a velocity built by the forward log law at a
    # known u_tau, along any wall-parallel direction, gives tau_wx =
    # u_tau^2 cos(theta); in the viscous sublayer u_tau^2 = nu |u| / h
    import oracle
    nu, kappa, B = 1e-4, 0.41, 5.2
    F, ref = [], []
    for ut, h, th in ((0.05, 0.1, 0.0), (0.03, 0.2, 0.7), (1.0, 0.05, -2.0), (0.2, 0.01, 3.0)):
        up = ut * (np.log(h * ut / nu) / kappa + B)
        F.append((up * np.cos(th), up * np.sin(th), h))
        ref.append(ut * ut * np.cos(th))
    h, up = 1e-3, 0.02              # y+ = h u_tau / nu = sqrt(h up / nu) = 0.45: linear law
    F.append((up, 0.0, h))
    ref.append(nu * up / h)
    got = om.equilibrium_alpha_2d(np.array(F), nu, kappa, B, oracle.loglaw_utau)
    assert np.allclose(got, ref, rtol=1e-13, atol=0)
