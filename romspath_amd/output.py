"""Trajectory output: dataOutput, createNetCDF and writeNetCDF
(ROMSPath.f90:1012-1072, 1140-2157) through the library's netCDF writer
(include/rp_ncout.h).

A run creates the file once with the release times (dob) and, whenever the
reference would call dataOutput (every iprint seconds of model time), samples
lon/lat on the GPU and appends one record of the particle columns in release
order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

FLAGS = {"WaterDepth": 1, "zeta": 2, "h": 4, "salttemp": 8, "bottom": 16, "collisions": 32}


def meta_of(prm, time_units="seconds since model start", run_name=None, behavior=0,
            executable=None, output_loc=None, administrator=None, institution=None,
            salt_temp_mean=None):
    """Global attributes from the namelist values (createNetCDF,
    ROMSPath.f90:1676-1836).  salt_temp_mean defaults to prm.SaltTempMean;
    the $output strings (ExeDir, OutDir, RunBy, Institution) are written
    empty when not given."""
    m = abi.RpNcMeta()
    enc = lambda s: s.encode() if s else None
    m.time_units = time_units.encode()
    m.run_name = enc(run_name)
    m.executable, m.output_loc = enc(executable), enc(output_loc)
    m.administrator, m.institution = enc(administrator), enc(institution)
    if salt_temp_mean is None:
        salt_temp_mean = getattr(prm, "SaltTempMean", False)
    m.salt_temp_mean = int(bool(salt_temp_mean))
    m.swimfast = float(getattr(prm, "swimfast", 0.0))
    m.temp_offset = float(getattr(prm, "TempOffset", 0.0))
    m.idt = int(prm.idt)
    m.days = float(prm.days)
    m.behavior = int(behavior)
    m.open_boundary = int(bool(prm.OpenOceanBoundary))
    m.vturb_on = int(bool(prm.VTurbOn))
    m.hturb_on = int(bool(prm.HTurbOn))
    m.seed = int(prm.seed) & 0x7FFFFFFF
    m.deltat, m.serr, m.smth, m.sub, m.aksback = (float(prm.deltat), float(prm.serr),
                                                   float(prm.smth), float(prm.sub),
                                                   float(prm.AKSback))
    return m


class TrajectoryFile:
    """One output file: createNetCDF(dob) on open, writeNetCDF per record."""

    def __init__(self, path, dob, ngrid, meta, extra=()):
        self._L = abi.lib()
        self.n = len(dob)
        self.ngrid = ngrid
        self.flags = 0
        for name in extra:
            self.flags |= FLAGS[name]
        self.h = C.c_void_p()
        d = np.ascontiguousarray(dob, dtype=np.float64)
        rc = self._L.rp_nc_create(C.byref(self.h), str(path).encode(), self.n, ngrid, self.flags,
                                  C.byref(meta), d.ctypes.data_as(abi.PD))
        if rc != 0:
            raise abi.RpError("rp_nc_create", rc, self._L.rp_nc_error(None).decode())

    def write(self, model_time, p, lon, lat):
        """p: the dict of rp_get_particles (release order), plus "salt" and "temp"
        (SaltTempOn) and "hob", "bustr", "bvstr" (WriteBottom) when the file
        has them; lon, lat likewise."""
        cols = [np.ascontiguousarray(p[k], dtype=np.float64) for k in ("age", "z", "status")]
        lo = np.ascontiguousarray(lon, dtype=np.float64)
        la = np.ascontiguousarray(lat, dtype=np.float64)
        gid = np.ascontiguousarray(p["gid"], dtype=np.int32)
        opt = {k: None if p.get(k) is None else np.ascontiguousarray(p[k], dtype=np.float64)
               for k in ("wd", "zeta", "bath", "salt", "temp", "hob", "bustr", "bvstr")}
        d = lambda a: None if a is None else a.ctypes.data_as(abi.PD)
        rc = self._L.rp_nc_write(self.h, float(model_time), d(cols[0]), d(lo), d(la), d(cols[1]),
                                 d(cols[2]), gid.ctypes.data_as(abi.PI32), d(opt["wd"]),
                                 d(opt["zeta"]), d(opt["bath"]), d(opt["salt"]), d(opt["temp"]),
                                 d(opt["hob"]), d(opt["bustr"]), d(opt["bvstr"]))
        if rc != 0:
            raise abi.RpError("rp_nc_write", rc, self._L.rp_nc_error(self.h).decode())

    @property
    def records(self):
        return int(self._L.rp_nc_records(self.h))

    def close(self):
        if self.h:
            rc = self._L.rp_nc_close(self.h)
            self.h = C.c_void_p()
            if rc != 0:
                raise abi.RpError("rp_nc_close", rc, "close failed")


def data_output(ctx, nc, model_time, lonlat=None):
    """dataOutput (ROMSPath.f90:1012-1025): lon/lat at every particle, then
    one record.  ctx: abi.Context after rp_set_lonlat for every grid (and
    rp_tracers_begin / rp_bottom_begin when the file has the SaltTempOn /
    WriteBottom variables: the means and P_HOB are read and reset,
    ROMSPath.f90:2009-2035, 2056-2065).  model_time is the record's
    model_time value, time + reftime (ROMSPath.f90:1931).  lonlat: (lon,
    lat) to write instead of sampling them, as the release record
    writeNetCDF(0, pLon, pLat) writes the release file's values
    (ROMSPath.f90:453)."""
    p = ctx.get_particles()
    if nc.flags & FLAGS["salttemp"]:
        p["salt"], p["temp"] = ctx.tracers_get(reset_mean=True)
    if nc.flags & FLAGS["bottom"]:
        p["hob"], p["bustr"], p["bvstr"] = ctx.bottom_get(reset_hob=True)
    lon, lat = ctx.sample_lonlat() if lonlat is None else lonlat
    nc.write(model_time, p, lon, lat)
    return p, lon, lat
