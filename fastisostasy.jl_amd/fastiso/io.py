"""Output path: NetcdfOutput and the output crops (src/io.jl:1-300).

Snapshots are read from the engine at the output times (fi_get_field), so only the fields a
snapshot names cross from device to host. The file is NetCDF-3 (64-bit offsets), written with
scipy.io.netcdf_file since the NetCDF C library is not part of this image. Variables, attributes
and dimension order follow NetCDF.jl as the reference calls it: a Julia variable over
[x, y, t] is stored as (t, y, x) in the file.
"""
import os

import numpy as np


class PaddedOutputCrop:
    """Drop pad cells on every side (src/io.jl:11-13,29-30)."""
    def __init__(self, pad=0):
        self.pad = int(pad)

    def bounds(self, nx, ny):
        return self.pad, nx - self.pad, self.pad, ny - self.pad


class AsymetricOutputCrop:
    """Drop pad_x1, pad_x2, pad_y1, pad_y2 cells (src/io.jl:20-25,33-39)."""
    def __init__(self, pad_x1=0, pad_x2=0, pad_y1=0, pad_y2=0):
        self.pads = (int(pad_x1), int(pad_x2), int(pad_y1), int(pad_y2))

    def bounds(self, nx, ny):
        a, b, c, d = self.pads
        return a, nx - b, c, ny - d


def crop(X, oc):
    """crop(x::Vector or X::Matrix, oc) (src/io.jl:29-39)."""
    X = np.asarray(X)
    if X.ndim == 1:
        n = X.shape[0]
        lo, hi, _, _ = oc.bounds(n, n)
        return X[lo:hi]
    i0, i1, j0, j1 = oc.bounds(*X.shape)
    return X[i0:i1, j0:j1]


def _ident(x):
    return x


# io_dict (src/io.jl:45-198): short name, long name, units and the map applied to parameters
IO_DICT = {
    "u": ("u", "Viscous displacement", "m", _ident),
    "ue": ("ue", "Elastic displacement", "m", _ident),
    "u_x": ("u_x", "Horizontal displacement in x", "m", _ident),
    "u_y": ("u_y", "Horizontal displacement in y", "m", _ident),
    "dudt": ("dudt", "Viscous displacement rate", "mm/yr", lambda x: 1e3 * x),
    "u_eq": ("u_eq", "Equilibrium displacement", "m", _ident),
    "H_ice": ("Hice", "Ice thickness", "m", _ident),
    "H_af": ("Haf", "Ice thickness above floatation", "m", _ident),
    "H_water": ("Hwater", "Water depth", "m", _ident),
    "z_b": ("z_b", "Bedrock position", "m", _ident),
    "z_ss": ("z_ss", "Sea-surface height (SSH)", "m", _ident),
    "dz_ss": ("dz_ss", "SSH perturbation", "m", _ident),
    "maskgrounded": ("maskgrounded", "Mask for grounded ice", "1", _ident),
    "maskocean": ("maskocean", "Mask for ocean", "1", _ident),
    "z_bsl": ("z_bsl", "Barystatic sea level", "m", _ident),
    "effective_viscosity": ("effective_viscosity", "Effective viscosity", "Pa s", np.log10),
    "tau": ("tau", "Relaxation time", "yr", _ident),
    "litho_thickness": ("litho_thickness", "Lithosphere thickness", "km", lambda x: 1e-3 * x),
    "pseudodiff_scaling": ("R", "Pseudo-differential scaling", "1", _ident),
    "maskactive": ("active_mask", "Mask where load is active", "1", _ident),
}


def _param(solidearth, name):
    v = getattr(solidearth, name)
    return np.asarray(v, dtype=np.float64)


class NetcdfOutput:
    """NetcdfOutput(domain, t, filename; vars3D, vars1D, params2D, Tout, output_crop, solidearth)
    (src/io.jl:208-292). The file is created here; write_nc (src/io.jl:302-319) adds snapshot k.
    As in the reference, only the parameters go through the io_dict map; a filename without
    '.nc' is ignored with a warning."""

    def __init__(self, domain, t, filename, vars3D=("u", "ue", "z_b", "dz_ss"), vars1D=("z_bsl",),
                 params2D=("effective_viscosity",), Tout=np.float32, output_crop=None, solidearth=None):
        self.t = [float(x) for x in t]
        self.filename = str(filename)
        self.vars3D, self.vars1D, self.params2D = list(vars3D), list(vars1D), list(params2D)
        self.Tout = np.dtype(Tout)
        self.oc = output_crop or PaddedOutputCrop(0)
        self.k = 1
        self.active = ".nc" in self.filename
        for v in self.vars3D + self.vars1D + self.params2D:
            if v not in IO_DICT:
                raise KeyError(f"no output variable {v!r}")
        if not self.active:
            import warnings
            warnings.warn("NetCDF filename does not end with '.nc' and is therefore ignored.")
            return
        from scipy.io import netcdf_file
        if os.path.isfile(self.filename):
            os.remove(self.filename)
        x, y = crop(domain.x, self.oc), crop(domain.y, self.oc)
        code = "f" if self.Tout == np.float32 else "d"
        with netcdf_file(self.filename, "w", version=2) as nc:
            for name, vals, long in (("x", x, "x"), ("y", y, "y"), ("t", self.t, "time")):
                nc.createDimension(name, len(vals))
                var = nc.createVariable(name, code, (name,))
                var[:] = np.asarray(vals, dtype=self.Tout)
                var.longname = long
                var.units = "yr" if name == "t" else "m"
            for v in self.vars3D:
                short, long, units, _ = IO_DICT[v]
                var = nc.createVariable(short, code, ("t", "y", "x"))
                var.longname, var.units = long, units
            for v in self.vars1D:
                short, long, units, _ = IO_DICT[v]
                var = nc.createVariable(short, code, ("t",))
                var.longname, var.units = long, units
            for v in self.params2D:
                short, long, units, fmap = IO_DICT[v]
                var = nc.createVariable(short, code, ("y", "x"))
                var.longname, var.units = long, units
                if solidearth is not None:
                    var[:, :] = np.asarray(fmap(crop(_param(solidearth, v), self.oc)), dtype=self.Tout).T

    def write(self, engine, k=None):
        """write_nc!(ncout, state, k) (src/io.jl:302-319): snapshot k (1-based) of vars3D and vars1D."""
        if not self.active:
            return
        from scipy.io import netcdf_file
        k = self.k if k is None else k
        stats = None
        with netcdf_file(self.filename, "a") as nc:
            for v in self.vars3D:
                X = crop(engine.get_field(v), self.oc)
                nc.variables[IO_DICT[v][0]][k - 1, :, :] = np.asarray(X, dtype=self.Tout).T
            for v in self.vars1D:
                stats = stats or engine.stats()
                nc.variables[IO_DICT[v][0]][k - 1] = self.Tout.type(stats[v])


def read_nc(filename, var):
    """A variable of a NetcdfOutput file in the reference's index order: (x, y, t), (x, y) or (t,)."""
    from scipy.io import netcdf_file
    with netcdf_file(filename, "r", mmap=False) as nc:
        a = np.array(nc.variables[var][:])
    return a.transpose(tuple(range(a.ndim))[::-1])
