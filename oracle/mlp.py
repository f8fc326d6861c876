"""CPU oracle of the learned wall-stress closure (test infrastructure only).

A numpy restatement, in FP64, of the network the library evaluates and trains
on the GPU: 4 -> 64 -> 64 -> 64 -> 2, SiLU, with BF16 rounding (round to
nearest even) applied at the same points as the GPU kernels when
`quant=True`:
  weights W1..W4 -> BF16; hidden activations h1, h2, h3 -> BF16 before the
  next product; in the backward pass dz3 and dz2 -> BF16 before the
  matrix-core products dH2 = W3 dZ3, dH1 = W2 dZ2, and the operands of the
  weight gradients dW2 = sum h1 dz2, dW3 = sum h2 dz3 -> BF16.
With `quant=False` every quantity is FP64, which is what the finite-
difference checks pin.  Loss: L = 1/(2n) sum over samples and both outputs
of (y - t)^2 on standardised outputs.  Adam: Kingma & Ba (2015), beta1 0.9,
beta2 0.999, eps 1e-8, bias-corrected.

The wall-cell features and equilibrium targets (`wall_samples`) restate
`wall_sample`/`wm_features` of the library with the log law of
oracle.c (`orc_loglaw_utau`).  There is no reference counterpart: the
reference's wall stress comes from its RL agent
(code/smarties_stat.f90:231-258); its wall-model point and spacing follow
code/wallmodel.f90:23-46.
"""
import numpy as np

NPARAM, NNORM = 8770, 12
W1, B1 = 0, 256
W2, B2 = 320, 4416
W3, B3 = 4480, 8576
W4, B4 = 8640, 8768


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(4, 64), p[B1:W2], p[W2:B2].reshape(64, 64), p[B2:W3],
            p[W3:B3].reshape(64, 64), p[B3:W4], p[W4:B4].reshape(64, 2), p[B4:NPARAM])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def standardise(X, norm):
    norm = np.asarray(norm, dtype=np.float64)
    return (np.asarray(X, np.float32).astype(np.float64) - norm[0:4]) / norm[4:8]


def forward(p, norm, X, quant=True):
    """Raw features (n, 4) -> raw outputs (n, 2)."""
    f = _forward(p, standardise(X, norm), quant)
    norm = np.asarray(norm, dtype=np.float64)
    return f["y"] * norm[10:12] + norm[8:10]


def _forward(p, x, quant):
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    z1 = x @ q(w1) + b1
    h1 = q(silu(z1))
    z2 = h1 @ q(w2) + b2
    h2 = q(silu(z2))
    z3 = h2 @ q(w3) + b3
    h3 = q(silu(z3))
    y = h3 @ q(w4) + b4
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, y=y)


def gradients(p, norm, X, T, quant=True):
    """(gradient (8770,), loss) of L = 1/(2n) sum (y - t)^2, standardised."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    norm = np.asarray(norm, dtype=np.float64)
    x = standardise(X, norm)
    t = (np.asarray(T, np.float32).astype(np.float64) - norm[8:10]) / norm[10:12]
    f = _forward(p, x, quant)
    n = len(x)
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    e = f["y"] - t
    loss = 0.5 * np.sum(e * e) / n
    dy = e / n
    g = np.zeros(NPARAM)
    g[W4:B4] = (f["h3"].T @ dy).ravel()
    g[B4:NPARAM] = dy.sum(0)
    dz3 = (dy @ q(w4).T) * dsilu(f["z3"])
    g[W3:B3] = (q(f["h2"]).T @ q(dz3)).ravel()
    g[B3:W4] = dz3.sum(0)
    dz2 = (q(dz3) @ q(w3).T) * dsilu(f["z2"])
    g[W2:B2] = (q(f["h1"]).T @ q(dz2)).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (q(dz2) @ q(w2).T) * dsilu(f["z1"])
    g[W1:B1] = (x.T @ dz1).ravel()
    g[B1:W2] = dz1.sum(0)
    return g, loss


def loss_only(p, norm, X, T, quant=False):
    norm = np.asarray(norm, dtype=np.float64)
    t = (np.asarray(T, np.float32).astype(np.float64) - norm[8:10]) / norm[10:12]
    y = _forward(p, standardise(X, norm), quant)["y"]
    return 0.5 * np.sum((y - t) ** 2) / len(t)


def adam_step(p, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step number t >= 1); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def fit_normalization(X, Y):
    X = np.asarray(X, np.float32).astype(np.float64)
    Y = np.asarray(Y, np.float32).astype(np.float64)
    sx, sy = X.std(0), Y.std(0)
    sx[~(sx > 1e-6)] = 1.0
    sy[~(sy > 1e-6)] = 1.0
    return np.concatenate([X.mean(0), sx, Y.mean(0), sy])


# ---------------------------------------------------------------- wall cells

def wall_samples(U, W, alpha, y, yg, nu, h_wm, kappa, B, dpdx, utau, flow=False):
    """Features (2*nx*nzg, 4) and equilibrium targets (., 2) of every wall cell
    (alpha_x order: i fastest, then wall, then k), FP64.  U (nx,nyg,nzg),
    W (nxg,nyg,nz), alpha (nx,2,nzg); y, yg 1-based arrays (index 0 unused);
    utau(u, h, nu, kappa, B) the log-law friction velocity.  flow=True also
    returns the sampled (u, w, h) of every cell (., 3)."""
    nx, nyg, nzg = U.shape
    ny = nyg - 1
    jb, wb, hb = 2, 1.0, yg[2] - y[1]
    if h_wm > 0.0:
        yt = y[1] + h_wm
        while jb < nyg - 2 and yg[jb + 1] < yt:
            jb += 1
        wb = (yg[jb + 1] - yt) / (yg[jb + 1] - yg[jb])
        hb = h_wm
    X = np.zeros((2 * nx * nzg, 4))
    Y = np.zeros((2 * nx * nzg, 2))
    F = np.zeros((2 * nx * nzg, 3))
    for k in range(1, nzg + 1):
        kc = nzg - 2 if k == 1 else (3 if k == nzg else k)
        for wall in range(2):
            ja = jb if wall == 0 else nyg + 1 - jb
            jn = jb + 1 if wall == 0 else nyg - jb
            for i in range(1, nx + 1):
                def uw(j):
                    u = U[i - 1, j - 1, k - 1]
                    w = 0.25 * (W[i - 1, j - 1, kc - 2] + W[i - 1, j - 1, kc - 1] + W[i, j - 1, kc - 2] +
                                W[i, j - 1, kc - 1])
                    return u, w
                ua, wa = uw(ja)
                un, wn = uw(jn)
                u = wb * ua + (1.0 - wb) * un
                w = wb * wa + (1.0 - wb) * wn
                dn = yg[jb + 1] - yg[jb] if wall == 0 else yg[ja] - yg[jn]
                dudy = (np.hypot(un, wn) - np.hypot(ua, wa)) / dn
                h = hb if (h_wm > 0.0 or wall == 0) else y[ny] - yg[nyg - 1]
                up = np.sqrt(u * u + w * w)
                c = (i - 1) + nx * (wall + 2 * (k - 1))
                X[c] = [np.arcsinh(h * up / nu / 10.0), np.arcsinh(h * h * dudy / nu / 10.0),
                        np.arcsinh(h * np.sqrt(abs(alpha[i - 1, wall, k - 1])) / nu / 10.0),
                        np.arcsinh(h ** 3 * dpdx / nu ** 2 / 10.0)]
                ut = utau(up, h, nu, kappa, B)
                Y[c] = [np.log1p(h * h * ut * ut / nu ** 2), 0.0]
                F[c] = [u, w, h]
    return (X, Y, F) if flow else (X, Y)


def equilibrium_alpha_2d(F, nu, kappa, B, utau):
    """The stress the wall model falls back to: tau_wx = u_tau(|u_par|)^2 u/|u_par|."""
    out = np.zeros(len(F))
    for c, (u, w, h) in enumerate(F):
        up = np.hypot(u, w)
        ut = utau(up, h, nu, kappa, B)
        out[c] = ut * ut * (u / up if up > 0 else 1.0)
    return out
