"""The reference's snapshot / .init file format.

Stream-unformatted, little-endian: for each of x, y, z, xm, ym, zm an int32
length then float64 data; then U (int32 nx,nyg,nzg + data), V (nxg,ny,nzg),
W (nxg,nyg,nz) and, in snapshots, P and nu_t (nxg,nyg,nzg), arrays in
Fortran order (code/input_output.f90:390-570 reader, :579-730 writer;
scripts/check_file.m:1-49).  Reading adds no noise (the reference adds
+-0.05 uniform noise on read, code/input_output.f90:541-563), so a restart is
exact.
"""
import numpy as np


def write(path, x, y, z, U, V, W, P=None, nu_t=None):
    x, y, z = (np.asarray(a, dtype="<f8") for a in (x, y, z))
    with open(path, "wb") as f:
        for a in (x, y, z, 0.5 * (x[:-1] + x[1:]), 0.5 * (y[:-1] + y[1:]), 0.5 * (z[:-1] + z[1:])):
            np.asarray([len(a)], dtype="<i4").tofile(f)
            np.asarray(a, dtype="<f8").tofile(f)
        for a in (U, V, W) + tuple(b for b in (P, nu_t) if b is not None):
            a = np.asarray(a, dtype="<f8")
            np.asarray(a.shape, dtype="<i4").tofile(f)
            a.ravel(order="F").tofile(f)


def read(path, with_extra=False):
    """Returns dict(x, y, z, xm, ym, zm, U, V, W[, P, nu_t])."""
    out = {}
    with open(path, "rb") as f:
        for name in ("x", "y", "z", "xm", "ym", "zm"):
            n = int(np.fromfile(f, dtype="<i4", count=1)[0])
            out[name] = np.fromfile(f, dtype="<f8", count=n)
        names = ["U", "V", "W"] + (["P", "nu_t"] if with_extra else [])
        for name in names:
            dims = np.fromfile(f, dtype="<i4", count=3)
            if len(dims) < 3:
                raise ValueError(f"{path}: missing {name}")
            n = int(np.prod(dims))
            data = np.fromfile(f, dtype="<f8", count=n)
            out[name] = data.reshape(tuple(int(d) for d in dims), order="F")
    nx, ny, nz = len(out["x"]), len(out["y"]), len(out["z"])
    exp = {"U": (nx, ny + 1, nz + 1), "V": (nx + 1, ny, nz + 1), "W": (nx + 1, ny + 1, nz),
           "P": (nx + 1, ny + 1, nz + 1), "nu_t": (nx + 1, ny + 1, nz + 1)}
    for name in names:
        if out[name].shape != exp[name]:
            raise ValueError(f"{path}: wrong size of {name}: {out[name].shape} != {exp[name]}")
    return out


def write_parameters(path, nx, ny, nz, CFL, nu, dPdx, x_mass_cte, LES_model, nsteps, nsave, nstats, nmonitor,
                     filein, fileout, nstep_init=0, random_init=0, Dirichlet_nu_t=1):
    """The reference's positional input file (scripts/case_folder/input_parameters_*,
    read by code/input_output.f90:20-91)."""
    with open(path, "w") as f:
        f.write(f"! domain size nx ny nz\n {nx} {ny} {nz}\n")
        f.write(f"! CFL (or fixed dT if negative)\n {CFL!r}\n")
        f.write(f"! viscosity\n  {nu!r}\n")
        f.write(f"! Pressure gradient dPdx, x_mass_cte (0->use dPdx, 1->constant mass flow)\n {dPdx!r} {x_mass_cte}\n")
        f.write(f"! LES model (0: no model 1: constant smagorinsky 2: DSM 3: AMD 4: Vreman)\n {LES_model}\n")
        f.write(f"! nsteps, nsave, nstats, nmonitor\n {nsteps}   {nsave}   {nstats}   {nmonitor}\n")
        f.write(f"! input flow field file\n '{filein}'\n")
        f.write(f"! output flow field file\n '{fileout}'\n")
        f.write(f"! first number for output flow field\n {nstep_init}\n")
        f.write(f"! random initial condition 0->no (reads previous file), 1->yes\n {random_init}\n")
        f.write(f"! (SGS) BC for eddy vicosity: 1->Dirichlet, 0->Neumann\n {Dirichlet_nu_t}\n")
