"""ROMS grid and history input: InitGrid (grid_module.f90:41-422) and
updateHydro (hydrodynamic_module.f90:112-760, getFileNames :774-808) over
the library's netCDF classic reader (include/rp_ncin.h).

* `NcFile`: one open file (dimensions, variable shapes, values as float64).
* `file_name`: getFileNames, the history file of number filenum.
* `read_grid`: the grid of one nest from its first history file, with
  reftime and time_units as InitGrid derives them from the time variable.
* `HydroFiles`: hydro records by (grid, record in file, file number), the
  fields updateHydro reads (zeta, u, v, w, AKs; salt and temp with
  SaltTempOn; bustr and bvstr with WriteBottom; the wet/dry masks of the
  -DWETDRY build; constants where a read flag is off), and the file roll-over
  the time schedule needs (tdim, multifile).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


class NcFile:
    """NF90_OPEN(path, NF90_NOWRITE) and the inquiries/reads updateHydro makes."""

    def __init__(self, path):
        self._L = abi.lib()
        self.path = str(path)
        self.h = C.c_void_p()
        rc = self._L.rp_ncin_open(C.byref(self.h), self.path.encode())
        if rc != 0:
            raise abi.RpError("rp_ncin_open", rc, self._L.rp_ncin_error(None).decode())

    def _chk(self, fn, rc):
        if rc != 0:
            raise abi.RpError(fn, rc, self._L.rp_ncin_error(self.h).decode())

    def dim(self, name):
        n = C.c_int64()
        self._chk("rp_ncin_dim", self._L.rp_ncin_dim(self.h, name.encode(), C.byref(n)))
        return n.value

    def var(self, name):
        """(shape in file order, is_record)."""
        nd, rec = C.c_int32(), C.c_int32()
        shape = (C.c_int64 * 8)()
        self._chk("rp_ncin_var", self._L.rp_ncin_var(self.h, name.encode(), C.byref(nd), shape,
                                                     C.byref(rec)))
        return tuple(shape[k] for k in range(nd.value)), bool(rec.value)

    def has(self, name):
        try:
            self.var(name)
            return True
        except abi.RpError:
            return False

    def read(self, name, rec=None):
        """Record rec (0-based) of a record variable, or all of a fixed one,
        as float64 in file order (numpy C order = the reference's Fortran
        array with its indices reversed)."""
        shape, is_rec = self.var(name)
        if is_rec:
            shape = shape[1:]
            if rec is None:
                raise ValueError(f"{name} is a record variable: give rec")
        out = np.empty(int(np.prod(shape)) if shape else 1)
        self._chk("rp_ncin_get", self._L.rp_ncin_get(self.h, name.encode(), -1 if rec is None else rec,
                                                     out.ctypes.data_as(abi.PD), out.size))
        return out.reshape(shape) if shape else out[0]

    def att_text(self, var, att):
        buf = C.create_string_buffer(4096)
        n = C.c_int64()
        self._chk("rp_ncin_att_text", self._L.rp_ncin_att_text(
            self.h, None if var is None else var.encode(), att.encode(), buf, 4096, C.byref(n)))
        return buf.value.decode("latin-1")

    def close(self):
        if self.h:
            self._L.rp_ncin_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def file_name(prefix, filenum, numdigits=4, suffix=".nc", multifile=True):
    """getFileNames (hydrodynamic_module.f90:774-808): prefix, filenum as
    Fortran I<d>.<d> (zero-padded; asterisks when it does not fit), suffix;
    just the prefix when multifile is off."""
    if not multifile:
        return prefix.strip()
    if not 1 <= numdigits <= 8:
        raise ValueError("numdigits must be 1..8 (getFileNames)")
    s = f"{int(filenum):0{numdigits}d}"
    if len(s) > numdigits:
        s = "*" * numdigits
    return prefix.strip() + s + suffix.strip()


def seconds_units(units):
    """InitGrid's conversion of the time variable's units (grid_module.f90:
    384-415): (factor to seconds, time_units of the output file).  The
    leading word is replaced by "seconds" at the fixed offsets the reference
    uses."""
    if "days" in units:
        return 86400.0, "seconds" + units[4:]
    if "hours" in units:
        return 3600.0, "seconds" + units[5:]
    if "minutes" in units:
        return 60.0, "seconds" + units[7:]
    if "seconds" in units:
        return 1.0, units
    raise ValueError(f"not a recognized time unit: {units!r}")


@dataclass
class RomsGrid:
    """GRIDS(ng) as InitGrid fills it; attribute names as the library's
    rp_grid wants them (numpy arrays in file order, i.e. (eta, xi))."""
    L: int
    M: int
    N: int
    vtransform: int
    vstretching: int
    theta_s: float
    theta_b: float
    hc: float
    spherical: int
    h: np.ndarray
    pm: np.ndarray
    pn: np.ndarray
    mask_rho: np.ndarray
    mask_u: np.ndarray
    mask_v: np.ndarray
    lon: np.ndarray          # lon_rho (spherical) or x_rho
    lat: np.ndarray          # lat_rho or y_rho
    angle: np.ndarray
    s_rho: np.ndarray
    cs_r: np.ndarray
    s_w: np.ndarray
    cs_w: np.ndarray
    reftime: float = 0.0     # first value of the time variable, in seconds
    time_units: str = "seconds"

    # the coordinates InitGrid keeps in lon_rho / lat_rho, which the nested
    # grids' offsets are matched on (grid_module.f90:437-481)
    @property
    def x_rho(self):
        return self.lon

    @property
    def y_rho(self):
        return self.lat


def _spherical(v):
    # ROMS writes spherical as an int flag; older files as the character T/F
    v = int(round(float(v)))
    return 1 if v in (1, ord("T"), ord("t")) else 0


def read_grid(path, time_vname="ocean_time"):
    """InitGrid for one nest (grid_module.f90:41-422) from its history file."""
    with NcFile(path) as f:
        L, M = f.dim("xi_rho"), f.dim("eta_rho")
        N, Nw = f.dim("s_rho"), f.dim("s_w")
        if (f.dim("xi_u"), f.dim("eta_u"), f.dim("xi_v"), f.dim("eta_v")) != (L - 1, M, L, M - 1):
            raise ValueError(f"{path}: staggered dimensions do not match xi_rho, eta_rho")
        if Nw != N + 1:
            raise ValueError(f"{path}: s_w must be s_rho + 1")
        def scalar(name):           # rank 0, or the first value as nf90_get_var into a scalar
            return float(np.ravel(f.read(name))[0])

        sph = _spherical(scalar("spherical"))
        lon, lat = ("lon_rho", "lat_rho") if sph else ("x_rho", "y_rho")
        t0 = f.read(time_vname, 0) if f.var(time_vname)[1] else scalar(time_vname)
        factor, units = seconds_units(f.att_text(time_vname, "units"))
        return RomsGrid(
            L=L, M=M, N=N,
            vtransform=int(scalar("Vtransform")), vstretching=int(scalar("Vstretching")),
            theta_s=scalar("theta_s"), theta_b=scalar("theta_b"), hc=scalar("hc"), spherical=sph,
            h=f.read("h"), pm=f.read("pm"), pn=f.read("pn"),
            mask_rho=f.read("mask_rho").astype(np.int32),
            mask_u=f.read("mask_u").astype(np.int32), mask_v=f.read("mask_v").astype(np.int32),
            lon=f.read(lon), lat=f.read(lat), angle=f.read("angle"),
            s_rho=f.read("s_rho"), cs_r=f.read("Cs_r"), s_w=f.read("s_w"), cs_w=f.read("Cs_w"),
            reftime=float(t0) * factor, time_units=units)


@dataclass
class ReadFlags:
    """The namelist's read*/const* pairs (hydroparam, ROMSPath.h) that
    updateHydro honours, TempOffset, and which optional fields to read."""
    readZeta: bool = True
    constZeta: float = 0.0
    readU: bool = True
    constU: float = 0.0
    readV: bool = True
    constV: float = 0.0
    readW: bool = True
    constW: float = 0.0
    readAks: bool = True
    constAks: float = 0.0
    readSalt: bool = False
    constSalt: float = 0.0
    readTemp: bool = False
    constTemp: float = 0.0
    TempOffset: float = 0.0
    SaltTempOn: bool = False
    WriteBottom: bool = False


class HydroFiles:
    """Hydro records read from history files, one prefix per nest.

    get(ng, tstep, filenum): updateHydro(.., tstep, tind) for nest ng
    (hydrodynamic_module.f90:112-760): record tstep (1-based) of file
    file_name(prefix[ng], filenum).  tdim(filenum) is the number of records of
    nest 1's file, which run_External_Timestep's roll-over compares tstep
    against (ROMSPath.f90:473-480)."""

    def __init__(self, prefixes, filenum=1, numdigits=4, suffix=".nc", multifile=False,
                 flags=None, time_dname="ocean_time", cache=True, stokes_prefixes=None,
                 wetdry=False):
        self.prefixes = list(prefixes)
        # -DWETDRY: wetdry_mask_rho / _u / _v of every record
        # (hydrodynamic_module.f90:510-570), for rp_set_hydro_wetdry
        self.wetdry = bool(wetdry)
        # -DSTOKES: ustokes / vstokes of the same record from a second file
        # series (stokesprefix, hydrodynamic_module.f90:573-660), handed on
        # as su / sv for the library to add to u / v
        self.stokes_prefixes = None if stokes_prefixes is None else list(stokes_prefixes)
        self.filenum0 = int(filenum)
        self.numdigits = numdigits
        self.suffix = suffix
        self.multifile = multifile
        self.flags = flags or ReadFlags()
        self.time_dname = time_dname
        self._files = {}
        # the last few records read (a record is read again only for the
        # repeated load of the schedule's third external step)
        self.cache = {} if cache else None
        self.cache_size = 6

    def path(self, ng, filenum, stokes=False):
        prefix = (self.stokes_prefixes if stokes else self.prefixes)[ng - 1]
        return file_name(prefix, filenum, self.numdigits, self.suffix, self.multifile)

    def _file(self, ng, filenum, stokes=False):
        key = (ng, filenum, stokes)
        if key not in self._files:
            # files of this series before the previous one are done with
            for k in [k for k in self._files if k[0] == ng and k[2] == stokes and k[1] < filenum - 1]:
                self._files.pop(k).close()
            self._files[key] = NcFile(self.path(ng, filenum, stokes))
        return self._files[key]

    def tdim(self, filenum):
        return self._file(1, filenum).dim(self.time_dname)

    def get(self, ng, tstep, filenum=None):
        filenum = self.filenum0 if filenum is None else filenum
        key = (ng, tstep, filenum)
        if self.cache is not None and key in self.cache:
            return self.cache[key]
        f = self._file(ng, filenum)
        fl = self.flags
        r = tstep - 1
        if not 0 <= r < f.dim(self.time_dname):
            raise IndexError(f"{f.path}: record {tstep} of {f.dim(self.time_dname)}")

        def fld(name, read, const, shape_from):
            if read:
                return f.read(name, r)
            return np.full(shape_from, float(const))

        L, M = f.dim("xi_rho"), f.dim("eta_rho")
        N, Nw = f.dim("s_rho"), f.dim("s_w")
        rec = {
            "zeta": fld("zeta", fl.readZeta, fl.constZeta, (M, L)),
            "u": fld("u", fl.readU, fl.constU, (N, M, L - 1)),
            "v": fld("v", fl.readV, fl.constV, (N, M - 1, L)),
            "w": fld("w", fl.readW, fl.constW, (Nw, M, L)),
            "aks": fld("AKs", fl.readAks, fl.constAks, (Nw, M, L)),
        }
        if fl.SaltTempOn:
            rec["salt"] = fld("salt", fl.readSalt, fl.constSalt, (N, M, L))
            rec["temp"] = (f.read("temp", r) + fl.TempOffset) if fl.readTemp else \
                np.full((N, M, L), float(fl.constTemp))
        if fl.WriteBottom:
            rec["bustr"] = f.read("bustr", r)
            rec["bvstr"] = f.read("bvstr", r)
        if self.wetdry:
            for name in ("wetdry_mask_rho", "wetdry_mask_u", "wetdry_mask_v"):
                rec[name] = f.read(name, r)
        if self.stokes_prefixes is not None:
            fs = self._file(ng, filenum, stokes=True)
            rec["su"] = fs.read("ustokes", r)
            rec["sv"] = fs.read("vstokes", r)
        if self.cache is not None:
            if len(self.cache) >= self.cache_size:
                self.cache.pop(next(iter(self.cache)))
            self.cache[key] = rec
        return rec

    def close(self):
        for f in self._files.values():
            f.close()
        self._files.clear()
