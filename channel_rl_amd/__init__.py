"""MI355X channel-flow LES time step (the hot path of wadaniel/channel_RL).

Python front of the C ABI in include/chanles.h (libchanles.so, HIP for
gfx950).  Method names follow the reference's Fortran subroutines; arrays are
numpy float64 in the reference's Fortran shapes (order='F').  The HIP library
is the only compute path: importing works without it, but creating a Channel
raises if libchanles.so is missing or no GPU is present.
"""
import ctypes
import os

import numpy as np

__all__ = ["Channel", "ChanError", "load_library", "LIB_PATH", "FP64", "FP32_BULK",
           "WALL_GIVEN", "WALL_EQUILIBRIUM"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libchanles.so")

FP64, FP32_BULK = 0, 1
WALL_GIVEN, WALL_EQUILIBRIUM = 0, 1  # WALL_MLP = 2 below
U, V, W, NU_T, P, ALPHA_X, RHS_U, RHS_V, RHS_W = range(9)

EXPORTS = [
    "chan_create", "chan_destroy", "chan_last_error", "chan_set_field", "chan_get_field",
    "chan_set_scalars", "chan_get_scalars", "chan_apply_boundary_conditions",
    "chan_compute_eddy_viscosity", "chan_compute_wall_model", "chan_compute_rhs",
    "chan_compute_projection_step", "chan_solve_poisson_equation",
    "chan_compute_mean_mass_flow_U", "chan_initialize_mass_flow",
    "chan_compute_dPx_for_constant_mass_flow", "chan_compute_dt",
    "chan_compute_time_step_RK3", "chan_check_divergence", "chan_set_wall_model",
    "chan_compute_equilibrium_wall_stress", "chan_plane_statistics", "chan_run_steps",
    "chan_synchronize", "chan_profile_steps", "chan_profile_count", "chan_profile_entry",
    "chan_stream_copy", "chan_tile_shape_count", "chan_set_tile_shape", "chan_get_tile_shape",
    "chan_autotune", "chan_mlp_create", "chan_mlp_destroy", "chan_mlp_last_error", "chan_mlp_get_params",
    "chan_mlp_set_params", "chan_mlp_fit_normalization", "chan_mlp_forward", "chan_mlp_gradients",
    "chan_mlp_train", "chan_wall_samples", "chan_set_wall_network", "chan_wall_model_stats", "chan_evaluate_wall_stress",
    "chan_ensemble_create", "chan_ensemble_destroy", "chan_ensemble_last_error", "chan_ensemble_size",
    "chan_ensemble_run_steps", "chan_ensemble_synchronize", "chan_wall_shear", "chan_rl_observe", "chan_rl_act",
    "chan_compute_pressure",
    "chan_compute_pressure_noalloc", "chan_ensemble_set_grouping", "chan_ensemble_group_count",
    "chan_set_wall_viscosity_limit",
]
KERNEL_RHS = 0
KERNEL_DSM = 1
WALL_MLP = 2
MLP_NPARAM, MLP_NNORM = 8770, 12


class ChanError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("x", ctypes.POINTER(ctypes.c_double)), ("y", ctypes.POINTER(ctypes.c_double)),
                ("z", ctypes.POINTER(ctypes.c_double)),
                ("nu", ctypes.c_double), ("CFL", ctypes.c_double), ("dPdx", ctypes.c_double),
                ("x_mass_cte", ctypes.c_int), ("LES_model", ctypes.c_int),
                ("Dirichlet_nu_t", ctypes.c_int), ("precision", ctypes.c_int), ("device", ctypes.c_int)]


_LIB = None
_D = ctypes.POINTER(ctypes.c_double)


def load_library(path=LIB_PATH):
    """Load libchanles.so (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ChanError(f"{path} is not built: run `python -m channel_rl_amd.build`")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.chan_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(vp)]
    L.chan_destroy.argtypes = [vp]
    L.chan_destroy.restype = None
    L.chan_last_error.argtypes = [vp]
    L.chan_last_error.restype = ctypes.c_char_p
    L.chan_set_field.argtypes = [vp, ctypes.c_int, _D]
    L.chan_get_field.argtypes = [vp, ctypes.c_int, _D]
    L.chan_set_scalars.argtypes = [vp] + [ctypes.c_double] * 4
    L.chan_get_scalars.argtypes = [vp, _D]
    for n in ("chan_apply_boundary_conditions", "chan_compute_eddy_viscosity", "chan_compute_wall_model",
              "chan_compute_rhs", "chan_compute_projection_step", "chan_solve_poisson_equation",
              "chan_initialize_mass_flow", "chan_compute_time_step_RK3", "chan_synchronize"):
        getattr(L, n).argtypes = [vp]
    L.chan_compute_mean_mass_flow_U.argtypes = [vp, _D]
    L.chan_compute_dPx_for_constant_mass_flow.argtypes = [vp, _D]
    L.chan_compute_dt.argtypes = [vp, _D]
    L.chan_check_divergence.argtypes = [vp, _D, _D]
    L.chan_set_wall_model.argtypes = [vp, ctypes.c_int] + [ctypes.c_double] * 3
    L.chan_compute_equilibrium_wall_stress.argtypes = [vp] + [ctypes.c_double] * 3
    L.chan_plane_statistics.argtypes = [vp, _D]
    L.chan_run_steps.argtypes = [vp, ctypes.c_int]
    L.chan_profile_steps.argtypes = [vp, ctypes.c_int]
    L.chan_profile_count.argtypes = [vp]
    L.chan_profile_entry.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _D,
                                     ctypes.POINTER(ctypes.c_long)]
    L.chan_stream_copy.argtypes = [vp, ctypes.c_long, ctypes.c_int, _D]
    L.chan_tile_shape_count.argtypes = [ctypes.c_int]
    L.chan_set_tile_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.chan_get_tile_shape.argtypes = [vp, ctypes.c_int]
    L.chan_autotune.argtypes = [vp, ctypes.c_int, _D, ctypes.c_int]
    _F = ctypes.POINTER(ctypes.c_float)
    L.chan_mlp_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.chan_mlp_destroy.argtypes = [vp]
    L.chan_mlp_last_error.argtypes = [vp]
    L.chan_mlp_last_error.restype = ctypes.c_char_p
    L.chan_mlp_get_params.argtypes = [vp, _F, _F]
    L.chan_mlp_set_params.argtypes = [vp, _F, _F]
    L.chan_mlp_fit_normalization.argtypes = [vp, _F, _F, ctypes.c_int]
    L.chan_mlp_forward.argtypes = [vp, _F, ctypes.c_int, _F]
    L.chan_mlp_gradients.argtypes = [vp, _F, _F, ctypes.c_int, _F, _D]
    L.chan_mlp_train.argtypes = [vp, _F, _F, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_ulonglong, _D]
    L.chan_wall_samples.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _F, _F,
                                    _D]
    L.chan_set_wall_network.argtypes = [vp, vp, ctypes.c_double]
    L.chan_wall_model_stats.argtypes = [vp, _D]
    L.chan_evaluate_wall_stress.argtypes = [vp]
    L.chan_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    L.chan_ensemble_destroy.argtypes = [vp]
    L.chan_ensemble_destroy.restype = None
    L.chan_ensemble_last_error.argtypes = [vp]
    L.chan_ensemble_last_error.restype = ctypes.c_char_p
    L.chan_ensemble_size.argtypes = [vp]
    L.chan_ensemble_run_steps.argtypes = [vp, ctypes.c_int]
    L.chan_ensemble_set_grouping.argtypes = [vp, ctypes.c_int]
    L.chan_set_wall_viscosity_limit.argtypes = [vp, ctypes.c_double]
    L.chan_ensemble_group_count.argtypes = [vp]
    L.chan_ensemble_synchronize.argtypes = [vp]
    L.chan_wall_shear.argtypes = [vp, _D]
    L.chan_rl_observe.argtypes = [vp, ctypes.c_int, ctypes.c_double, _D, _D, _D]
    L.chan_rl_act.argtypes = [vp, ctypes.c_int, _D]
    L.chan_compute_pressure.argtypes = [vp]
    L.chan_compute_pressure_noalloc.argtypes = [vp]
    _LIB = L
    return L


def _p(a):
    return a.ctypes.data_as(_D)


def _pf(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected an (n, {cols}) array, got {a.shape}")
    return a


class Ensemble:
    """Channels advanced together, one graph launch per ensemble step."""

    def __init__(self, channels):
        L = load_library()
        self.channels = list(channels)  # the members must outlive the ensemble
        arr = (ctypes.c_void_p * len(self.channels))(*[c._h.value for c in self.channels])
        h = ctypes.c_void_p()
        rc = L.chan_ensemble_create(arr, len(self.channels), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = L.chan_ensemble_last_error(h).decode() if h.value else "chan_ensemble_create failed"
            L.chan_ensemble_destroy(h)
            self._h = None
            raise ChanError(f"chan_ensemble_create: {msg} (status {rc})")

    def close(self):
        if getattr(self, "_h", None):
            load_library().chan_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            raise ChanError(f"{what}: {load_library().chan_ensemble_last_error(self._h).decode()} (status {rc})")

    def __len__(self):
        return load_library().chan_ensemble_size(self._h)

    def run_steps(self, n):
        self._chk(load_library().chan_ensemble_run_steps(self._h, int(n)), "run_steps")

    def set_grouping(self, on):
        """Grouped launches for members of one shape (on, the default), or
        every member's own kernels (off)."""
        self._chk(load_library().chan_ensemble_set_grouping(self._h, 1 if on else 0), "set_grouping")

    def group_count(self):
        """Groups and lone members the members' current settings give."""
        n = load_library().chan_ensemble_group_count(self._h)
        if n < 0:
            self._chk(n, "group_count")
        return n

    def synchronize(self):
        self._chk(load_library().chan_ensemble_synchronize(self._h), "synchronize")


class WallNetwork:
    """The learned wall-stress closure (a chan_mlp): 4 -> 64 -> 64 -> 64 -> 2,
    SiLU, BF16 weights with FP32 accumulation, trained with Adam on the GPU.
    Features X are (n, 4) and outputs Y (n, 2), raw (unstandardised)."""

    def __init__(self, seed=0, device=0):
        L = load_library()
        h = ctypes.c_void_p()
        rc = L.chan_mlp_create(device, seed, ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = L.chan_mlp_last_error(h).decode() if h.value else "chan_mlp_create failed"
            L.chan_mlp_destroy(h)
            self._h = None
            raise ChanError(f"chan_mlp_create: {msg} (status {rc})")

    def close(self):
        """Release the network; refused (ChanError) while a channel still uses
        it as its wall model."""
        if getattr(self, "_h", None):
            self._chk(load_library().chan_mlp_destroy(self._h), "close")
            self._h = None

    def __del__(self):
        # channels hold a reference while attached, so none is left here
        try:
            self.close()
        except ChanError:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise ChanError(f"{what}: {load_library().chan_mlp_last_error(self._h).decode()} (status {rc})")

    def params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        n = np.zeros(MLP_NNORM, np.float32)
        self._chk(load_library().chan_mlp_get_params(self._h, _pf(p), _pf(n)), "get_params")
        return p, n

    def set_params(self, params=None, norm=None):
        pp = None if params is None else np.ascontiguousarray(params, np.float32)
        nn = None if norm is None else np.ascontiguousarray(norm, np.float32)
        if pp is not None and pp.shape != (MLP_NPARAM,):
            raise ValueError("params must hold 8770 values")
        if nn is not None and nn.shape != (MLP_NNORM,):
            raise ValueError("norm must hold 12 values")
        self._chk(load_library().chan_mlp_set_params(self._h, None if pp is None else _pf(pp),
                                                     None if nn is None else _pf(nn)), "set_params")

    def fit_normalization(self, X, Y):
        X, Y = _f32(X, 4), _f32(Y, 2)
        self._chk(load_library().chan_mlp_fit_normalization(self._h, _pf(X), _pf(Y), len(X)), "fit_normalization")

    def forward(self, X):
        X = _f32(X, 4)
        Y = np.zeros((len(X), 2), np.float32)
        self._chk(load_library().chan_mlp_forward(self._h, _pf(X), len(X), _pf(Y)), "forward")
        return Y

    def gradients(self, X, Y):
        X, Y = _f32(X, 4), _f32(Y, 2)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = ctypes.c_double()
        self._chk(load_library().chan_mlp_gradients(self._h, _pf(X), _pf(Y), len(X), _pf(g), ctypes.byref(loss)),
                  "gradients")
        return g, loss.value

    def train(self, X, Y, epochs=20, batch=1024, lr=1e-3, seed=0):
        X, Y = _f32(X, 4), _f32(Y, 2)
        losses = np.zeros(max(epochs, 1))
        self._chk(load_library().chan_mlp_train(self._h, _pf(X), _pf(Y), len(X), int(epochs), int(batch),
                                                float(lr), int(seed), _p(losses)), "train")
        return losses[:epochs]


class Channel:
    """One channel resident on the GPU (a chan_ctx)."""

    def __init__(self, x, y, z, nu, CFL=-1e-2, dPdx=0.0, x_mass_cte=0, LES_model=2,
                 Dirichlet_nu_t=1, precision=FP64, device=0):
        L = load_library()
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        self.nx, self.ny, self.nz = (len(a) for a in self._keep)
        self.nxg, self.nyg, self.nzg = self.nx + 1, self.ny + 1, self.nz + 1
        self.precision = precision
        self.nu = nu
        cfg = _Config(self.nx, self.ny, self.nz, _p(self._keep[0]), _p(self._keep[1]), _p(self._keep[2]),
                      nu, CFL, dPdx, x_mass_cte, LES_model, Dirichlet_nu_t, precision, device)
        h = ctypes.c_void_p()
        rc = L.chan_create(ctypes.byref(cfg), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = L.chan_last_error(h).decode() if h.value else "chan_create failed"
            L.chan_destroy(h)
            self._h = None
            raise ChanError(f"chan_create: {msg} (status {rc})")

    def close(self):
        if getattr(self, "_h", None):
            load_library().chan_destroy(self._h)
            self._h = None
        self._net = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            raise ChanError(f"{what}: {load_library().chan_last_error(self._h).decode()} (status {rc})")

    def shape(self, which):
        nx, ny, nz, nxg, nyg, nzg = self.nx, self.ny, self.nz, self.nxg, self.nyg, self.nzg
        return {U: (nx, nyg, nzg), V: (nxg, ny, nzg), W: (nxg, nyg, nz), NU_T: (nxg, nyg, nzg),
                P: (nxg, nyg, nzg), ALPHA_X: (nx, 2, nzg), RHS_U: (nx - 2, nyg - 2, nzg - 2),
                RHS_V: (nxg - 2, ny - 2, nzg - 2), RHS_W: (nxg - 2, nyg - 2, nz - 2)}[which]

    def set_field(self, which, a):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape(which):
            raise ValueError(f"field {which}: shape {a.shape} != {self.shape(which)}")
        self._chk(load_library().chan_set_field(self._h, which, _p(a)), "chan_set_field")

    def get_field(self, which):
        a = np.zeros(self.shape(which), dtype=np.float64, order="F")
        self._chk(load_library().chan_get_field(self._h, which, _p(a)), "chan_get_field")
        return a

    def set_state(self, U_, V_, W_):
        self.set_field(U, U_); self.set_field(V, V_); self.set_field(W, W_)

    def state(self):
        return self.get_field(U), self.get_field(V), self.get_field(W)

    def set_scalars(self, dt=0.0, t=0.0, dPdx=0.0, Qflow_x_0=0.0):
        self._chk(load_library().chan_set_scalars(self._h, dt, t, dPdx, Qflow_x_0), "chan_set_scalars")

    def scalars(self):
        out = np.zeros(4)
        self._chk(load_library().chan_get_scalars(self._h, _p(out)), "chan_get_scalars")
        return dict(dt=out[0], t=out[1], dPdx=out[2], Qflow_x_0=out[3])

    # the reference's subroutines
    def apply_boundary_conditions(self):
        self._chk(load_library().chan_apply_boundary_conditions(self._h), "apply_boundary_conditions")

    def compute_eddy_viscosity(self):
        self._chk(load_library().chan_compute_eddy_viscosity(self._h), "compute_eddy_viscosity")

    def compute_wall_model(self):
        self._chk(load_library().chan_compute_wall_model(self._h), "compute_wall_model")

    def compute_rhs(self):
        self._chk(load_library().chan_compute_rhs(self._h), "compute_rhs")
        return self.get_field(RHS_U), self.get_field(RHS_V), self.get_field(RHS_W)

    def compute_projection_step(self):
        self._chk(load_library().chan_compute_projection_step(self._h), "compute_projection_step")

    def solve_poisson_equation(self):
        self._chk(load_library().chan_solve_poisson_equation(self._h), "solve_poisson_equation")

    def compute_mean_mass_flow_U(self):
        q = ctypes.c_double()
        self._chk(load_library().chan_compute_mean_mass_flow_U(self._h, ctypes.byref(q)), "compute_mean_mass_flow_U")
        return q.value

    def initialize_mass_flow(self):
        self._chk(load_library().chan_initialize_mass_flow(self._h), "initialize_mass_flow")

    def compute_dPx_for_constant_mass_flow(self):
        q = ctypes.c_double()
        self._chk(load_library().chan_compute_dPx_for_constant_mass_flow(self._h, ctypes.byref(q)),
                  "compute_dPx_for_constant_mass_flow")
        return q.value

    def compute_dt(self, sync=True):
        q = ctypes.c_double()
        self._chk(load_library().chan_compute_dt(self._h, ctypes.byref(q) if sync else None), "compute_dt")
        return q.value if sync else None

    def compute_time_step_RK3(self):
        self._chk(load_library().chan_compute_time_step_RK3(self._h), "compute_time_step_RK3")

    def check_divergence(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(load_library().chan_check_divergence(self._h, ctypes.byref(a), ctypes.byref(b)), "check_divergence")
        return a.value, b.value

    def set_wall_model(self, mode, h_wm=0.0, kappa=0.41, B=5.2):
        self._chk(load_library().chan_set_wall_model(self._h, mode, h_wm, kappa, B), "set_wall_model")

    def set_wall_viscosity_limit(self, limit):
        """Hold the wall model's total viscosity in [0, limit dy^2/dt] (0: off)."""
        self._chk(load_library().chan_set_wall_viscosity_limit(self._h, float(limit)), "set_wall_viscosity_limit")

    def compute_equilibrium_wall_stress(self, h_wm=0.0, kappa=0.41, B=5.2):
        self._chk(load_library().chan_compute_equilibrium_wall_stress(self._h, h_wm, kappa, B),
                  "compute_equilibrium_wall_stress")

    def plane_statistics(self):
        out = np.zeros(7 * self.ny)
        self._chk(load_library().chan_plane_statistics(self._h, _p(out)), "plane_statistics")
        return out.reshape(7, self.ny)

    def run_steps(self, n):
        self._chk(load_library().chan_run_steps(self._h, n), "run_steps")

    def profile_steps(self, n):
        """Per-kernel HIP-event timing of n un-captured steps:
        {kernel: (total_ms, launches)}."""
        L = load_library()
        self._chk(L.chan_profile_steps(self._h, n), "profile_steps")
        out = {}
        buf = ctypes.create_string_buffer(128)
        for q in range(L.chan_profile_count(self._h)):
            ms, cnt = ctypes.c_double(), ctypes.c_long()
            self._chk(L.chan_profile_entry(self._h, q, buf, 128, ctypes.byref(ms), ctypes.byref(cnt)),
                      "profile_entry")
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out

    def stream_copy(self, nbytes, reps=20):
        """Achieved bandwidth of a plain device copy: (GB/s counting read + write, ms per copy)."""
        ms = ctypes.c_double()
        self._chk(load_library().chan_stream_copy(self._h, int(nbytes), int(reps), ctypes.byref(ms)), "stream_copy")
        return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value

    def synchronize(self):
        self._chk(load_library().chan_synchronize(self._h), "synchronize")

    def wall_samples(self, h_wm=0.0, nu=None, kappa=0.41, B=5.2, stress=False):
        """Features (2*nx*nzg, 4) and equilibrium targets (2*nx*nzg, 2) of every
        wall cell at height h_wm, with viscosity nu (default: the channel's);
        stress=True also returns the equilibrium stress u_tau^2 per cell."""
        n = 2 * self.nx * self.nzg
        X = np.zeros((n, 4), np.float32)
        Y = np.zeros((n, 2), np.float32)
        T = np.zeros(n)
        nu = self.nu if nu is None else nu
        self._chk(load_library().chan_wall_samples(self._h, h_wm, nu, kappa, B, _pf(X), _pf(Y), _p(T)),
                  "wall_samples")
        return (X, Y, T) if stress else (X, Y)

    def set_wall_network(self, net, max_rel_err=0.05):
        self._net = net  # keep it alive while attached
        self._chk(load_library().chan_set_wall_network(self._h, None if net is None else net._h, max_rel_err),
                  "set_wall_network")

    def wall_shear(self):
        """dict(tau_bottom, tau_top, u_tau, Re_tau) of the current state."""
        out = np.zeros(4)
        self._chk(load_library().chan_wall_shear(self._h, _p(out)), "wall_shear")
        return dict(tau_bottom=out[0], tau_top=out[1], u_tau=out[2], Re_tau=out[3])

    def compute_pressure(self):
        """The true pressure of the current state (code/pressure.f90:21)."""
        self._chk(load_library().chan_compute_pressure(self._h), "compute_pressure")
        return self.get_field(P)

    def rl_observe(self, nagents, dPdx_ref, draws):
        """States (2, nagents, 6) and rewards (2, nagents) of the RL agents;
        draws (2, nagents, 2) uniform numbers."""
        d = np.ascontiguousarray(draws, dtype=np.float64).reshape(2 * nagents * 2)
        st = np.zeros(2 * nagents * 6)
        rw = np.zeros(2 * nagents)
        self._chk(load_library().chan_rl_observe(self._h, nagents, dPdx_ref, _p(d), _p(st), _p(rw)), "rl_observe")
        return st.reshape(2, nagents, 6), rw.reshape(2, nagents)

    def rl_act(self, nagents, factors):
        f = np.ascontiguousarray(factors, dtype=np.float64).reshape(2 * nagents)
        self._chk(load_library().chan_rl_act(self._h, nagents, _p(f)), "rl_act")

    def evaluate_wall_stress(self):
        """Evaluate the selected wall-stress source once into alpha_x."""
        self._chk(load_library().chan_evaluate_wall_stress(self._h), "evaluate_wall_stress")

    def wall_model_stats(self):
        """(mean relative deviation from the equilibrium stress, fallback fraction)
        of the last learned-wall-model evaluation."""
        out = np.zeros(2)
        self._chk(load_library().chan_wall_model_stats(self._h, _p(out)), "wall_model_stats")
        return float(out[0]), float(out[1])

    def tile_shape(self, kernel=KERNEL_RHS):
        return load_library().chan_get_tile_shape(self._h, kernel)

    def set_tile_shape(self, shape, kernel=KERNEL_RHS):
        self._chk(load_library().chan_set_tile_shape(self._h, kernel, int(shape)), "set_tile_shape")

    def autotune(self, reps=10):
        """Time every tile shape of the RHS and the DSM pass-A kernels on this
        channel and keep the fastest of each; returns {"rhs": ms per launch of
        each RHS shape, "dsm": the same for the DSM shapes}."""
        L = load_library()
        nr, nd = L.chan_tile_shape_count(KERNEL_RHS), L.chan_tile_shape_count(KERNEL_DSM)
        t = np.zeros(nr + nd)
        self._chk(L.chan_autotune(self._h, int(reps), _p(t), nr + nd), "autotune")
        return {"rhs": t[:nr], "dsm": t[nr:]}
