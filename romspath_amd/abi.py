"""ctypes binding of librp.so (include/rp_abi.h).

The library must be built (romspath_amd/build.py); there is no fallback:
importing this module without the library, or creating a context on a host
without a GPU, raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "librp.so"
MAXGRID = 5

RP_PREC_FP32_FIELDS = 0
RP_PREC_FP64 = 1

ERRORS = {-1: "RP_E_ARG", -2: "RP_E_STATE", -3: "RP_E_HIP", -4: "RP_E_ALLOC",
          -5: "RP_E_RANGE", -6: "RP_E_NODEV", -7: "RP_E_IO"}

PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)
PI32 = C.POINTER(C.c_int32)
PU8 = C.POINTER(C.c_uint8)
PI64 = C.POINTER(C.c_int64)

EXPORTS = ["rp_create", "rp_set_grid", "rp_set_hydro_slot", "rp_set_particles", "rp_step",
           "rp_step_noalloc",
           "rp_get_particles", "rp_set_deviates", "rp_sample_deviates", "rp_active_steps",
           "rp_secant_capped", "rp_synchronize", "rp_kernel_timing", "rp_phase_timing",
           "rp_modal_basis",
           "rp_set_tuning", "rp_set_scenarios", "rp_step_scenario", "rp_step_batch",
           "rp_set_perturbation", "rp_adjoint_record", "rp_adjoint",
           "rp_set_lonlat", "rp_sample_lonlat", "rp_set_grid_angle", "rp_locate_particles",
           "rp_set_hydro_tracers", "rp_sample_tracers",
           "rp_set_hydro_bstress", "rp_set_hydro_wetdry", "rp_bottom_begin", "rp_bottom_update", "rp_bottom_get",
           "rp_tracers_begin", "rp_tracers_update", "rp_tracers_get",
           "rp_nc_create", "rp_nc_write", "rp_nc_records", "rp_nc_error", "rp_nc_close",
           "rp_ncin_open", "rp_ncin_dim", "rp_ncin_var", "rp_ncin_get", "rp_ncin_att_text",
           "rp_ncin_error", "rp_ncin_close",
           "rp_last_error", "rp_destroy"]
# entry points whose result is not an int status
_RESTYPE = {"rp_last_error": C.c_char_p, "rp_nc_error": C.c_char_p, "rp_nc_records": C.c_int64,
            "rp_ncin_error": C.c_char_p}


class RpNcMeta(C.Structure):
    """rp_nc_meta of include/rp_ncout.h: the file's global attributes."""
    _fields_ = [("time_units", C.c_char_p), ("run_name", C.c_char_p), ("idt", C.c_int32),
                ("days", C.c_float), ("behavior", C.c_int32), ("open_boundary", C.c_int32),
                ("vturb_on", C.c_int32), ("hturb_on", C.c_int32), ("seed", C.c_int32),
                ("deltat", C.c_double), ("serr", C.c_double), ("smth", C.c_double),
                ("sub", C.c_double), ("aksback", C.c_double), ("swimfast", C.c_double),
                ("temp_offset", C.c_double), ("salt_temp_mean", C.c_int32),
                ("executable", C.c_char_p), ("output_loc", C.c_char_p),
                ("administrator", C.c_char_p), ("institution", C.c_char_p)]


class RpConfig(C.Structure):
    _fields_ = [("idt", C.c_int32), ("dt", C.c_int32), ("deltat", C.c_double),
                ("serr", C.c_double), ("smth", C.c_double), ("sub", C.c_double),
                ("aksback", C.c_double), ("vturb_on", C.c_int32), ("hturb_on", C.c_int32),
                ("kh", C.c_double * MAXGRID), ("scheme", C.c_int32), ("nsb", C.c_int32),
                ("vertdist", C.c_double), ("open_boundary", C.c_int32), ("ngrid", C.c_int32),
                ("seed", C.c_uint64), ("precision", C.c_int32), ("block", C.c_int32),
                ("sort_interval", C.c_int32), ("reserved", C.c_int32 * 5)]


class RpGrid(C.Structure):
    _fields_ = [("L", C.c_int32), ("M", C.c_int32), ("N", C.c_int32), ("vtransform", C.c_int32),
                ("hc", C.c_double), ("s_rho", PD), ("cs_r", PD), ("s_w", PD), ("cs_w", PD),
                ("h", PD), ("pm", PD), ("pn", PD), ("mask_rho", PI32), ("mask_u", PI32),
                ("mask_v", PI32), ("scl", (C.c_double * 2) * MAXGRID),
                ("off", (C.c_double * 2) * MAXGRID)]


class RpError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"librp.so not built: {LIB_PATH} (python -m romspath_amd.build)")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.rp_create.argtypes = [C.POINTER(vp), C.POINTER(RpConfig)]
        L.rp_set_grid.argtypes = [vp, C.c_int32, C.POINTER(RpGrid)]
        L.rp_set_hydro_slot.argtypes = [vp, C.c_int32, C.c_int32, PD, PD, PD, PD, PD, PD, PD]
        L.rp_set_particles.argtypes = [vp, C.c_int64, PD, PD, PD, PD, PD, PD, PI32, PU8]
        L.rp_step.argtypes = [vp, PD, PD, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                              C.c_int32]
        L.rp_step_noalloc.argtypes = [vp, PD, PD, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int64, C.c_int32]
        L.rp_get_particles.argtypes = [vp, C.c_int64, PD, PD, PD, PD, PD, PI32, PU8, PD, PD, PD]
        L.rp_set_deviates.argtypes = [vp, PD, C.c_int64, C.c_int32]
        L.rp_sample_deviates.argtypes = [vp, C.c_int64, C.c_int64, C.c_int32, PD]
        L.rp_active_steps.argtypes = [vp, PI64]
        L.rp_secant_capped.argtypes = [vp, PI64]
        L.rp_synchronize.argtypes = [vp]
        L.rp_kernel_timing.argtypes = [vp, PD, PI64]
        L.rp_phase_timing.argtypes = [vp, PD, PD]
        L.rp_modal_basis.argtypes = [C.c_int32, PF, PD]
        L.rp_set_tuning.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.rp_set_scenarios.argtypes = [vp, C.c_int32, PI64]
        L.rp_step_scenario.argtypes = [vp, C.c_int32, PD, PD, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int64, C.c_int32]
        L.rp_step_batch.argtypes = [vp, C.c_int32, PD, PD, PI32, PI32, PI64, PI32, C.c_int32]
        L.rp_set_perturbation.argtypes = [vp, C.c_int32, C.c_int32, PD]
        L.rp_adjoint_record.argtypes = [vp, C.c_int32]
        L.rp_adjoint.argtypes = [vp, C.c_int64, PD, PD, PD, PD, PD, PD, PD]
        L.rp_set_lonlat.argtypes = [vp, C.c_int32, PD, PD]
        L.rp_set_grid_angle.argtypes = [vp, C.c_int32, C.c_int32, PD]
        L.rp_set_hydro_tracers.argtypes = [vp, C.c_int32, C.c_int32, PD, PD]
        L.rp_sample_tracers.argtypes = [vp, C.c_int64, PD, PD, C.c_int32, C.c_int32, C.c_int32, PD, PD]
        L.rp_tracers_begin.argtypes = [vp, C.c_int32, C.c_int32]
        L.rp_tracers_update.argtypes = [vp, PD, PD, C.c_int32, C.c_int32, C.c_int32]
        L.rp_tracers_get.argtypes = [vp, C.c_int64, PD, PD, C.c_int32]
        L.rp_set_hydro_bstress.argtypes = [vp, C.c_int32, C.c_int32, PD, PD]
        L.rp_set_hydro_wetdry.argtypes = [vp, C.c_int32, C.c_int32, PD, PD, PD]
        L.rp_bottom_begin.argtypes = [vp]
        L.rp_bottom_update.argtypes = [vp, PD, PD, C.c_int32, C.c_int32, C.c_int32]
        L.rp_bottom_get.argtypes = [vp, C.c_int64, PD, PD, PD, C.c_int32]
        L.rp_locate_particles.argtypes = [vp, C.c_int64, PD, PD, PD, C.c_int32, PD, PD, PD, PI32, PU8]
        L.rp_sample_lonlat.argtypes = [vp, C.c_int64, PD, PD]
        L.rp_nc_create.argtypes = [C.POINTER(vp), C.c_char_p, C.c_int64, C.c_int32, C.c_int32,
                                   C.POINTER(RpNcMeta), PD]
        L.rp_ncin_open.argtypes = [C.POINTER(vp), C.c_char_p]
        L.rp_ncin_dim.argtypes = [vp, C.c_char_p, PI64]
        L.rp_ncin_var.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int32), PI64, C.POINTER(C.c_int32)]
        L.rp_ncin_get.argtypes = [vp, C.c_char_p, C.c_int64, PD, C.c_int64]
        L.rp_ncin_att_text.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64, PI64]
        L.rp_ncin_error.argtypes = [vp]
        L.rp_ncin_close.argtypes = [vp]
        L.rp_nc_write.argtypes = [vp, C.c_double, PD, PD, PD, PD, PD, PI32, PD, PD, PD, PD, PD, PD, PD, PD]
        L.rp_nc_records.argtypes = [vp]
        L.rp_nc_error.argtypes = [vp]
        L.rp_nc_close.argtypes = [vp]
        L.rp_last_error.argtypes = [vp]
        L.rp_destroy.argtypes = [vp]
        for name in EXPORTS:
            getattr(L, name).restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(PD)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_config(prm, ngrid=1, precision="fp32", block=0, sort_interval=0) -> RpConfig:
    c = RpConfig()
    c.idt = prm.idt
    c.dt = prm.dt
    c.deltat = prm.deltat
    c.serr = prm.serr
    c.smth = prm.smth
    c.sub = prm.sub
    c.aksback = prm.AKSback
    c.vturb_on = int(prm.VTurbOn)
    c.hturb_on = int(prm.HTurbOn)
    for i in range(MAXGRID):
        c.kh[i] = prm.ConstantHTurb[i] if i < len(prm.ConstantHTurb) else 0.0
    c.scheme = prm.scheme
    c.nsb = prm.nsb
    c.vertdist = prm.vertdist
    c.open_boundary = int(prm.OpenOceanBoundary)
    c.ngrid = ngrid
    c.seed = prm.seed
    c.precision = RP_PREC_FP64 if precision == "fp64" else RP_PREC_FP32_FIELDS
    c.block = block
    c.sort_interval = sort_interval
    return c


class Context:
    """One librp context (one HIP stream)."""

    def __init__(self, cfg: RpConfig):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.rp_create(C.byref(h), C.byref(cfg))
        if rc != 0:
            raise RpError("rp_create", rc, "cannot create context")
        self.h = h
        self.cfg = cfg
        self.n = 0
        self._keep = []

    def _chk(self, fn, rc):
        if rc != 0:
            msg = self._L.rp_last_error(self.h).decode(errors="replace")
            raise RpError(fn, rc, msg)

    def close(self):
        if self.h:
            self._L.rp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, ng, grid, scl=None, off=None):
        g = RpGrid()
        g.L, g.M, g.N = grid.L, grid.M, grid.N
        g.vtransform = grid.vtransform
        g.hc = grid.hc
        keep = {}
        for name in ("s_rho", "cs_r", "s_w", "cs_w", "h", "pm", "pn"):
            keep[name] = _f64(getattr(grid, name))
            setattr(g, name, keep[name].ctypes.data_as(PD))
        for name in ("mask_rho", "mask_u", "mask_v"):
            keep[name] = np.ascontiguousarray(getattr(grid, name), dtype=np.int32)
            setattr(g, name, keep[name].ctypes.data_as(PI32))
        for i in range(MAXGRID):
            for k in range(2):
                g.scl[i][k] = 1.0 if scl is None else scl[i][k]
                g.off[i][k] = 0.0 if off is None else off[i][k]
        self._chk("rp_set_grid", self._L.rp_set_grid(self.h, ng, C.byref(g)))

    def set_hydro_slot(self, ng, slot, rec):
        a = {k: _f64(rec[k]) for k in ("zeta", "u", "v", "w", "aks")}
        su = _f64(rec["su"]) if "su" in rec else None
        sv = _f64(rec["sv"]) if "sv" in rec else None
        self._chk("rp_set_hydro_slot", self._L.rp_set_hydro_slot(
            self.h, ng, slot, _d(a["zeta"]), _d(a["u"]), _d(a["v"]), _d(a["w"]), _d(a["aks"]),
            _d(su), _d(sv)))

    def set_particles(self, parts):
        n = parts.n
        cols = [_f64(parts.x), _f64(parts.y), _f64(parts.z), _f64(parts.age), _f64(parts.dob),
                _f64(parts.status)]
        gid = np.ascontiguousarray(parts.gid, dtype=np.int32)
        isin = np.ascontiguousarray(parts.isin, dtype=np.uint8)
        self._chk("rp_set_particles", self._L.rp_set_particles(
            self.h, n, *[_d(c) for c in cols], gid.ctypes.data_as(PI32),
            isin.ctypes.data_as(PU8)))
        self.n = n

    def step(self, ex, ix, ets, tb, tc, tf, step_index, want_diag=False):
        exa = (C.c_double * 3)(*ex)
        ixa = (C.c_double * 3)(*ix)
        self._chk("rp_step", self._L.rp_step(self.h, exa, ixa, ets, tb, tc, tf, step_index,
                                             int(want_diag)))

    def step_noalloc(self, ex, ix, ets, tb, tc, tf, step_index, want_diag=False):
        exa = (C.c_double * 3)(*ex)
        ixa = (C.c_double * 3)(*ix)
        self._chk("rp_step_noalloc", self._L.rp_step_noalloc(self.h, exa, ixa, ets, tb, tc, tf,
                                                             step_index, int(want_diag)))

    def set_tuning(self, block=0, sort_interval=0, tile=(1, 1)):
        self._chk("rp_set_tuning", self._L.rp_set_tuning(self.h, int(block), int(sort_interval),
                                                         int(tile[0]), int(tile[1])))

    def set_scenarios(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.int64)
        self._chk("rp_set_scenarios", self._L.rp_set_scenarios(
            self.h, int(c.size), c.ctypes.data_as(PI64)))

    def step_scenario(self, scen, ex, ix, ets, tb, tc, tf, step_index, want_diag=False):
        exa = (C.c_double * 3)(*ex)
        ixa = (C.c_double * 3)(*ix)
        self._chk("rp_step_scenario", self._L.rp_step_scenario(
            self.h, int(scen), exa, ixa, ets, tb, tc, tf, step_index, int(want_diag)))

    def step_batch(self, steps, use_graph=True):
        """steps: sequence of objects with ex, ix, ets, tb, tc, tf, step_index, output."""
        k = len(steps)
        ex = np.ascontiguousarray([st.ex for st in steps], dtype=np.float64).reshape(k, 3)
        ix = np.ascontiguousarray([st.ix for st in steps], dtype=np.float64).reshape(k, 3)
        ets = np.ascontiguousarray([st.ets for st in steps], dtype=np.int32)
        slots = np.ascontiguousarray([(st.tb, st.tc, st.tf) for st in steps], dtype=np.int32)
        sidx = np.ascontiguousarray([st.step_index for st in steps], dtype=np.int64)
        diag = np.ascontiguousarray([int(getattr(st, "output", False)) for st in steps],
                                    dtype=np.int32)
        self._chk("rp_step_batch", self._L.rp_step_batch(
            self.h, k, _d(ex), _d(ix), ets.ctypes.data_as(PI32), slots.ctypes.data_as(PI32),
            sidx.ctypes.data_as(PI64), diag.ctypes.data_as(PI32), int(bool(use_graph))))

    def set_perturbation(self, coef):
        """coef: array [2][ny][nx] (du, dv in m/s) or None."""
        if coef is None:
            self._chk("rp_set_perturbation", self._L.rp_set_perturbation(self.h, 0, 0, None))
            return
        c = _f64(coef)
        self._chk("rp_set_perturbation", self._L.rp_set_perturbation(
            self.h, int(c.shape[2]), int(c.shape[1]), _d(c)))

    def adjoint_record(self, ckpt_every):
        self._chk("rp_adjoint_record", self._L.rp_adjoint_record(self.h, int(ckpt_every)))

    def adjoint(self, lam, want_pert=False, pert_shape=None):
        """lam: dict or tuple of dL/dx, dL/dy, dL/dz (release order).  Returns
        (gx, gy, gz) and, with want_pert, dL/dcoef shaped pert_shape."""
        lx, ly, lz = (_f64(a) for a in lam)
        n = self.n
        gx, gy, gz = np.empty(n), np.empty(n), np.empty(n)
        gp = np.empty(pert_shape) if want_pert else None
        self._chk("rp_adjoint", self._L.rp_adjoint(self.h, n, _d(lx), _d(ly), _d(lz), _d(gx), _d(gy),
                                                   _d(gz), _d(gp)))
        return (gx, gy, gz), gp

    def get_particles(self):
        n = self.n
        out = {k: np.empty(n) for k in ("x", "y", "z", "age", "status", "wd", "zeta", "bath")}
        gid = np.empty(n, dtype=np.int32)
        isin = np.empty(n, dtype=np.uint8)
        self._chk("rp_get_particles", self._L.rp_get_particles(
            self.h, n, _d(out["x"]), _d(out["y"]), _d(out["z"]), _d(out["age"]),
            _d(out["status"]), gid.ctypes.data_as(PI32), isin.ctypes.data_as(PU8),
            _d(out["wd"]), _d(out["zeta"]), _d(out["bath"])))
        out["gid"] = gid
        out["isin"] = isin
        return out

    def set_deviates(self, normals):
        if normals is None:
            self._chk("rp_set_deviates", self._L.rp_set_deviates(self.h, None, 0, 0))
            return
        a = _f64(normals)
        self._chk("rp_set_deviates", self._L.rp_set_deviates(self.h, _d(a), a.shape[0], a.shape[1]))

    def set_lonlat(self, ng, lon, lat):
        lo, la = _f64(lon), _f64(lat)
        self._chk("rp_set_lonlat", self._L.rp_set_lonlat(self.h, ng, _d(lo), _d(la)))

    def set_hydro_tracers(self, ng, slot, salt, temp):
        sa, te = _f64(salt), _f64(temp)
        self._chk("rp_set_hydro_tracers", self._L.rp_set_hydro_tracers(self.h, ng, slot, _d(sa), _d(te)))

    def sample_tracers(self, ex, ix, tb, tc, tf):
        salt, temp = np.empty(self.n), np.empty(self.n)
        exa = (C.c_double * 3)(*ex)
        ixa = (C.c_double * 3)(*ix)
        self._chk("rp_sample_tracers", self._L.rp_sample_tracers(
            self.h, self.n, exa, ixa, tb, tc, tf, _d(salt), _d(temp)))
        return salt, temp

    def tracers_begin(self, mean, t_c):
        self._chk("rp_tracers_begin", self._L.rp_tracers_begin(self.h, int(bool(mean)), t_c))

    def tracers_update(self, ex, ix, tb, tc, tf):
        exa = (C.c_double * 3)(*ex)
        ixa = (C.c_double * 3)(*ix)
        self._chk("rp_tracers_update", self._L.rp_tracers_update(self.h, exa, ixa, tb, tc, tf))

    def tracers_get(self, reset_mean=False):
        """P_Salt, P_Temp (or their means since the last reset) in release order."""
        salt, temp = np.empty(self.n), np.empty(self.n)
        self._chk("rp_tracers_get", self._L.rp_tracers_get(self.h, self.n, _d(salt), _d(temp),
                                                           int(bool(reset_mean))))
        return salt, temp

    def set_hydro_bstress(self, ng, slot, bustr, bvstr):
        bu, bv = _f64(bustr), _f64(bvstr)
        self._chk("rp_set_hydro_bstress", self._L.rp_set_hydro_bstress(self.h, ng, slot, _d(bu), _d(bv)))

    def set_hydro_wetdry(self, ng, slot, mask_rho, mask_u, mask_v):
        """The reference built with -DWETDRY: the record's wetdry_mask_rho/u/v
        (0/1) into slot 1..3."""
        mr, mu, mv = _f64(mask_rho), _f64(mask_u), _f64(mask_v)
        self._chk("rp_set_hydro_wetdry",
                  self._L.rp_set_hydro_wetdry(self.h, ng, slot, _d(mr), _d(mu), _d(mv)))

    def bottom_begin(self):
        self._chk("rp_bottom_begin", self._L.rp_bottom_begin(self.h))

    def bottom_update(self, ex, ix, tb, tc, tf):
        exa = (C.c_double * 3)(*ex)
        ixa = (C.c_double * 3)(*ix)
        self._chk("rp_bottom_update", self._L.rp_bottom_update(self.h, exa, ixa, tb, tc, tf))

    def bottom_get(self, reset_hob=False):
        """P_HOB, P_bustr, P_bvstr in release order."""
        hob, bu, bv = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        self._chk("rp_bottom_get", self._L.rp_bottom_get(self.h, self.n, _d(hob), _d(bu), _d(bv),
                                                         int(bool(reset_hob))))
        return hob, bu, bv

    def set_grid_angle(self, ng, spherical, angle=None):
        a = None if angle is None else _f64(angle)
        self._chk("rp_set_grid_angle", self._L.rp_set_grid_angle(self.h, ng, int(spherical), _d(a)))

    def locate_particles(self, lon, lat, z, t_b=1):
        """ini_ROMSPath's placement: returns dict x, y, z, status, gid, isin."""
        lo, la = _f64(lon), _f64(lat)
        n = lo.size
        out = {"x": np.empty(n), "y": np.empty(n), "z": np.array(z, dtype=np.float64),
               "status": np.empty(n), "gid": np.empty(n, dtype=np.int32),
               "isin": np.empty(n, dtype=np.uint8)}
        self._chk("rp_locate_particles", self._L.rp_locate_particles(
            self.h, n, _d(lo), _d(la), _d(out["z"]), int(t_b), _d(out["x"]), _d(out["y"]),
            _d(out["status"]), out["gid"].ctypes.data_as(PI32), out["isin"].ctypes.data_as(PU8)))
        return out

    def sample_lonlat(self):
        lon, lat = np.empty(self.n), np.empty(self.n)
        self._chk("rp_sample_lonlat", self._L.rp_sample_lonlat(self.h, self.n, _d(lon), _d(lat)))
        return lon, lat

    def sample_deviates(self, step_index, n, ndev):
        out = np.empty((n, ndev))
        self._chk("rp_sample_deviates",
                  self._L.rp_sample_deviates(self.h, step_index, n, ndev, _d(out)))
        return out

    def active_steps(self):
        v = C.c_int64()
        self._chk("rp_active_steps", self._L.rp_active_steps(self.h, C.byref(v)))
        return v.value

    def secant_capped(self):
        v = C.c_int64()
        self._chk("rp_secant_capped", self._L.rp_secant_capped(self.h, C.byref(v)))
        return v.value

    def synchronize(self):
        self._chk("rp_synchronize", self._L.rp_synchronize(self.h))

    def kernel_timing(self):
        ms = C.c_double()
        k = C.c_int64()
        self._chk("rp_kernel_timing", self._L.rp_kernel_timing(self.h, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def phase_timing(self):
        """(advection ms, vertical-walk ms) averages of the last kernel_timing()."""
        a = C.c_double()
        b = C.c_double()
        self._chk("rp_phase_timing", self._L.rp_phase_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value
