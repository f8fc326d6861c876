"""Running statistics of a channel (code/statistics.f90:22-145): plane means
of U, V, W, their squares and UV taken on the GPU every few steps, the wall
shear and u_tau, and the profiles in wall units.  Used to compare runs (the
FP32 bulk against FP64, a wall model against another) over many steps."""
import numpy as np

from . import Channel


class ProfileAverager:
    """Accumulates chan_plane_statistics and chan_wall_shear samples."""

    def __init__(self, ch: Channel, w0=None, w1=None):
        self.ch = ch
        self.n = 0
        self.sum = np.zeros((7, ch.ny))
        self.tau = np.zeros(2)
        # weighted interpolation of centre profiles to the y faces
        # (code/initialization.f90:398-400 weight_y_0/1), from the grid
        y = np.asarray(ch._keep[1])
        yg = np.concatenate([[2 * y[0] - 0.5 * (y[0] + y[1])], 0.5 * (y[:-1] + y[1:]),
                             [2 * y[-1] - 0.5 * (y[-2] + y[-1])]])
        self.y, self.yg = y, yg
        self.w1 = (y - yg[:-1]) / (yg[1:] - yg[:-1])
        self.w0 = 1.0 - self.w1

    def add(self):
        self.sum += self.ch.plane_statistics()
        w = self.ch.wall_shear()
        self.tau += (w["tau_bottom"], w["tau_top"])
        self.n += 1

    def run(self, steps, every):
        """Advance `steps` steps, sampling every `every` steps."""
        for _ in range(steps // every):
            self.ch.run_steps(every)
            self.add()

    def result(self):
        """Means over the samples: dict with y (faces), yc (centres), U, V, W,
        uu, vv, ww (centred second moments), uv (<u'v'> at the faces), u_tau,
        Re_tau, and the wall-unit profiles Uplus (centres) and uvplus.  Row q
        of the centre quantities is centre row q+1 (q = 0 the bottom ghost
        row), row q of V and uv is y face q+1."""
        m = self.sum / max(self.n, 1)
        U, V, W, U2, V2, W2, UV = m
        tau = self.tau / max(self.n, 1)
        Ly = self.y[-1] - self.y[0]
        ut = np.sqrt((tau[0] + tau[1]) / Ly)
        Uf = self.w0 * U + self.w1 * np.concatenate([U[1:], U[-1:]])
        uv = UV - Uf * V
        return dict(y=self.y, yc=self.yg[:-1], U=U, V=V, W=W, uu=U2 - U * U, vv=V2 - V * V, ww=W2 - W * W, uv=uv,
                    u_tau=ut, Re_tau=ut * Ly / 2.0 / self.ch.nu, Uplus=U / ut, uvplus=uv / (ut * ut), samples=self.n)


def profile_difference(a, b):
    """Relative differences of two results of ProfileAverager: the mean U+
    profile (interior centre rows, relative to max U+) and <u'v'>+ (interior
    faces, relative to max |<u'v'>+|)."""
    ua, ub = a["Uplus"][1:], b["Uplus"][1:]
    va, vb = a["uvplus"][1:-1], b["uvplus"][1:-1]
    return dict(Uplus=float(np.max(np.abs(ua - ub)) / np.max(np.abs(ua))),
                uvplus=float(np.max(np.abs(va - vb)) / np.max(np.abs(va))),
                u_tau=float(abs(a["u_tau"] - b["u_tau"]) / a["u_tau"]))
