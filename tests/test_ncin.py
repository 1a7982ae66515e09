"""netCDF classic input (include/rp_ncin.h, romspath_amd/roms.py): files
written by scipy's independent netCDF writer (CDF-1 and CDF-2) and a CDF-5
file built byte by byte from the format specification read back value for
value; InitGrid's grid, reftime and time_units; getFileNames; the
run_External_Timestep file roll-over; and an oracle run whose records come
from a set of history files equals the run on the same records in memory,
bit for bit.  No GPU.
"""
import struct

import numpy as np
import pytest
from scipy.io import netcdf_file

from oracle import oracle as O
from romspath_amd import abi, synth
from romspath_amd.roms import HydroFiles, NcFile, file_name, read_grid, seconds_units
from romspath_amd.runner import RecordSource
from romspath_amd.schedule import schedule


def write_history(path, g, recs, times, version=2, units="seconds since 2000-01-01 00:00:00",
                  ftype="d"):
    """A ROMS-like history file: grid variables and one record per entry of
    recs (dicts of zeta, u, v, w, aks and optional salt, temp, bustr, bvstr and
    the -DWETDRY masks wetdry_mask_rho/_u/_v)."""
    L, M, N = g.L, g.M, g.N
    f = netcdf_file(str(path), "w", version=version)
    for name, n in (("ocean_time", None), ("xi_rho", L), ("eta_rho", M), ("xi_u", L - 1),
                    ("eta_u", M), ("xi_v", L), ("eta_v", M - 1), ("s_rho", N), ("s_w", N + 1)):
        f.createDimension(name, n)

    def put(name, dims, data, t="d"):
        v = f.createVariable(name, t, dims)
        v[...] = data
        return v

    # scalars as one-element variables: scipy's writer places rank-0 variables
    # inside the record area of a file that has record variables
    f.createDimension("one", 1)
    put("spherical", ("one",), 0, "i")
    put("Vtransform", ("one",), g.vtransform, "i")
    put("Vstretching", ("one",), 4, "i")
    put("theta_s", ("one",), g.theta_s)
    put("theta_b", ("one",), g.theta_b)
    put("hc", ("one",), g.hc)
    rho = ("eta_rho", "xi_rho")
    for name, a in (("h", g.h), ("pm", g.pm), ("pn", g.pn), ("angle", np.zeros((M, L))),
                    ("x_rho", g.x_rho), ("y_rho", g.y_rho), ("mask_rho", g.mask_rho)):
        put(name, rho, a)
    put("mask_u", ("eta_u", "xi_u"), g.mask_u)
    put("mask_v", ("eta_v", "xi_v"), g.mask_v)
    put("s_rho", ("s_rho",), g.s_rho)
    put("Cs_r", ("s_rho",), g.cs_r)
    put("s_w", ("s_w",), g.s_w)
    put("Cs_w", ("s_w",), g.cs_w)
    t = f.createVariable("ocean_time", "d", ("ocean_time",))
    t.units = units
    dims = {"zeta": ("ocean_time",) + rho, "u": ("ocean_time", "s_rho", "eta_u", "xi_u"),
            "v": ("ocean_time", "s_rho", "eta_v", "xi_v"), "w": ("ocean_time", "s_w") + rho,
            "AKs": ("ocean_time", "s_w") + rho, "salt": ("ocean_time", "s_rho") + rho,
            "temp": ("ocean_time", "s_rho") + rho, "bustr": ("ocean_time", "eta_u", "xi_u"),
            "bvstr": ("ocean_time", "eta_v", "xi_v"), "wetdry_mask_rho": ("ocean_time",) + rho,
            "wetdry_mask_u": ("ocean_time", "eta_u", "xi_u"), "wetdry_mask_v": ("ocean_time", "eta_v", "xi_v")}
    keys = [k for k in ("zeta", "u", "v", "w", "aks", "salt", "temp", "bustr", "bvstr", "wetdry_mask_rho",
                        "wetdry_mask_u", "wetdry_mask_v") if k in recs[0]]
    vars_ = {k: f.createVariable("AKs" if k == "aks" else k, ftype, dims["AKs" if k == "aks" else k])
             for k in keys}
    for r, (rec, tt) in enumerate(zip(recs, times)):
        t[r] = tt
        for k in keys:
            vars_[k][r] = rec[k]
    f.close()


@pytest.mark.parametrize("version", [1, 2])
def test_reads_what_scipy_wrote(tmp_path, version):
    rng = np.random.default_rng(version)
    path = tmp_path / "a.nc"
    f = netcdf_file(str(path), "w", version=version)
    f.createDimension("t", None)       # scipy wants the unlimited dimension first
    f.createDimension("x", 5)
    f.createDimension("y", 3)
    f.title = "probe"
    data = {}
    for name, t in (("b", "b"), ("h", "h"), ("i", "i"), ("f", "f"), ("d", "d")):
        v = f.createVariable("fix_" + name, t, ("y", "x"))
        a = rng.integers(-100, 100, (3, 5)).astype(t) if t in "bhi" else rng.standard_normal((3, 5)).astype(t)
        v[...] = a
        data["fix_" + name] = a
        r = f.createVariable("rec_" + name, t, ("t", "x"))
        r.units = "m s-1"
        a = rng.integers(-100, 100, (4, 5)).astype(t) if t in "bhi" else rng.standard_normal((4, 5)).astype(t)
        r[:] = a
        data["rec_" + name] = a
    f.close()
    with NcFile(path) as g:
        assert g.dim("x") == 5 and g.dim("t") == 4
        assert g.var("rec_f") == ((4, 5), True) and g.var("fix_h") == ((3, 5), False)
        assert g.att_text(None, "title") == "probe" and g.att_text("rec_d", "units") == "m s-1"
        for name, a in data.items():
            if name.startswith("fix_"):
                assert np.array_equal(g.read(name), a.astype(np.float64)), name
            else:
                for r in range(4):
                    assert np.array_equal(g.read(name, r), a[r].astype(np.float64)), (name, r)


@pytest.mark.parametrize("t", ["h", "b"])
def test_single_record_variable_unpadded(tmp_path, t):
    # one record variable of a 2- or 1-byte type: records are not padded to 4 bytes
    path = tmp_path / "s.nc"
    f = netcdf_file(str(path), "w", version=1)
    f.createDimension("t", None)
    f.createDimension("x", 3)
    v = f.createVariable("r", t, ("t", "x"))
    a = np.arange(15).reshape(5, 3).astype(t) - 7
    v[:] = a
    f.close()
    ref = netcdf_file(str(path), "r", mmap=False)
    want = ref.variables["r"][:].copy()
    ref.close()
    with NcFile(path) as g:
        got = np.array([g.read("r", r) for r in range(g.dim("t"))])
    assert np.array_equal(got, want.astype(np.float64))


def _cdf5_bytes():
    """A CDF-5 file by the format specification: dims x = 3, t unlimited;
    fixed int64 a(x), uint16 b(x) and a rank-0 double s; record double
    r(t, x), two records."""
    be = lambda fmt, *v: struct.pack(">" + fmt, *v)

    def name(s):
        b = s.encode()
        return be("Q", len(b)) + b + b"\0" * ((4 - len(b) % 4) % 4)

    hdr = b"CDF\x05" + be("Q", 2)
    hdr += be("I", 0x0A) + be("Q", 2) + name("x") + be("Q", 3) + name("t") + be("Q", 0)
    hdr += be("I", 0) + be("Q", 0)                       # no global attributes
    hdr += be("I", 0x0B) + be("Q", 4)
    vars_ = [("a", [0], 10, 24), ("b", [0], 8, 8), ("s", [], 6, 8), ("r", [1, 0], 6, 24)]
    sizes = []
    for nm, dims, typ, vsize in vars_:
        sizes.append(name(nm) + be("Q", len(dims)) + b"".join(be("Q", d) for d in dims)
                     + be("I", 0) + be("Q", 0) + be("I", typ) + be("Q", vsize))
    head_len = len(hdr) + sum(len(s) + 8 for s in sizes)
    begin_a = head_len
    begin_b = begin_a + 24
    begin_s = begin_b + 8
    begin_r = begin_s + 8
    body = b""
    for s, beg in zip(sizes, (begin_a, begin_b, begin_s, begin_r)):
        body += s + be("Q", beg)
    data = be("3q", -5, 2**40, 7) + be("3H", 1, 65535, 3) + b"\0\0" + be("d", 3.25)
    data += be("3d", 0.5, -1.5, 2.25) + be("3d", 1e300, -0.0, 3.0)
    return hdr + body + data


def test_cdf5(tmp_path):
    path = tmp_path / "c5.nc"
    path.write_bytes(_cdf5_bytes())
    with NcFile(path) as g:
        assert g.dim("t") == 2 and g.var("r") == ((2, 3), True)
        assert np.array_equal(g.read("a"), [-5.0, 2.0**40, 7.0])
        assert np.array_equal(g.read("b"), [1.0, 65535.0, 3.0])
        assert g.read("s") == 3.25 and g.var("s") == ((), False)
        assert np.array_equal(g.read("r", 0), [0.5, -1.5, 2.25])
        assert np.array_equal(g.read("r", 1), [1e300, -0.0, 3.0])


def test_errors(tmp_path):
    h5 = tmp_path / "h5.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 100)
    with pytest.raises(abi.RpError) as e:
        NcFile(h5)
    assert e.value.code == -7 and "HDF5" in str(e.value)
    bad = tmp_path / "bad.nc"
    bad.write_bytes(_cdf5_bytes()[:40])
    with pytest.raises(abi.RpError) as e:
        NcFile(bad)
    assert e.value.code == -7
    with pytest.raises(abi.RpError) as e:
        NcFile(tmp_path / "missing.nc")
    assert e.value.code == -7
    ok = tmp_path / "c5.nc"
    ok.write_bytes(_cdf5_bytes())
    with NcFile(ok) as g:
        with pytest.raises(abi.RpError) as e:
            g.read("nope")
        assert e.value.code == -1
        with pytest.raises(abi.RpError) as e:
            g.read("r", 2)
        assert e.value.code == -5                      # RP_E_RANGE
        with pytest.raises(abi.RpError) as e:
            g.dim("y")
        assert e.value.code == -1


def test_file_names_and_time_units():
    assert file_name("ocean_his_", 7, 4, ".nc") == "ocean_his_0007.nc"
    assert file_name("his", 123, 2, ".nc") == "his**.nc"          # Fortran I2.2 overflow
    assert file_name("/d/one.nc ", 9, 4, ".nc", multifile=False) == "/d/one.nc"
    assert seconds_units("days since 2000-01-01 00:00:00") == (86400.0, "seconds since 2000-01-01 00:00:00")
    assert seconds_units("hours since 1990-1-1") == (3600.0, "seconds since 1990-1-1")
    assert seconds_units("minutes since 1990-1-1") == (60.0, "seconds since 1990-1-1")
    assert seconds_units("seconds since 1970-01-01") == (1.0, "seconds since 1970-01-01")
    with pytest.raises(ValueError):
        seconds_units("fortnights")


class _Files:
    def __init__(self, tdim, multifile=True, filenum0=3):
        self.t, self.multifile, self.filenum0 = tdim, multifile, filenum0

    def tdim(self, filenum):
        return self.t


def test_schedule_rolls_over_files():
    # dt = idt: one internal step per external step; 12 external steps
    plain = [ld for st in schedule(3600, 3600, 0.5, 3600) for ld in st.loads]
    rolled = [ld for st in schedule(3600, 3600, 0.5, 3600, hydro=_Files(4)) for ld in st.loads]
    assert [ld.slot for ld in plain] == [ld.slot for ld in rolled]
    # tstep moves to record 1 of the next file once it passes 4
    assert [(ld.filenum, ld.record) for ld in rolled] == [
        (3, 1), (3, 2), (3, 3), (3, 3), (3, 4), (4, 1), (4, 2), (4, 3), (4, 4), (5, 1), (5, 2),
        (5, 3), (5, 4)]
    assert [ld.record for ld in plain] == [(f - 3) * 4 + r for f, r in
                                          ((ld.filenum, ld.record) for ld in rolled)]
    with pytest.raises(RuntimeError):
        list(schedule(3600, 3600, 0.5, 3600, hydro=_Files(4, multifile=False)))


def history_set(tmp_path, cfg, tdim, nfiles, ftype="d"):
    """cfg's records 1 .. tdim*nfiles in files his_0001.nc, ...; returns the prefix."""
    recs = RecordSource(cfg)
    g = cfg.grids[0]
    for k in range(nfiles):
        ids = range(k * tdim + 1, (k + 1) * tdim + 1)
        write_history(tmp_path / f"his_{k + 1:04d}.nc", g, [recs.get(1, r) for r in ids],
                      [(r - 1) * cfg.params.dt for r in ids], ftype=ftype)
    return str(tmp_path / "his_")


def test_grid_and_records_from_files(tmp_path):
    cfg = synth.config_mini(n_particles=10)
    prefix = history_set(tmp_path, cfg, 4, 2)
    g, rg = cfg.grids[0], read_grid(prefix + "0001.nc")
    assert (rg.L, rg.M, rg.N, rg.vtransform, rg.spherical) == (g.L, g.M, g.N, g.vtransform, 0)
    for k in ("h", "pm", "pn", "mask_rho", "mask_u", "mask_v", "s_rho", "cs_r", "s_w", "cs_w"):
        assert np.array_equal(getattr(rg, k), getattr(g, k)), k
    assert np.array_equal(rg.lon, g.x_rho) and rg.reftime == 0.0
    assert rg.time_units == "seconds since 2000-01-01 00:00:00"
    hf = HydroFiles([prefix], filenum=1, multifile=True)
    recs = RecordSource(cfg)
    for filenum, tstep, glob in ((1, 1, 1), (1, 4, 4), (2, 1, 5), (2, 3, 7)):
        a, b = hf.get(1, tstep, filenum), recs.get(1, glob)
        for k in ("zeta", "u", "v", "w", "aks"):
            assert np.array_equal(a[k], b[k]), (filenum, tstep, k)
    assert hf.tdim(1) == 4
    hf.close()


def test_oracle_run_from_files_equals_in_memory(tmp_path):
    cfg = synth.config_mini(n_particles=300, days=0.5)
    prefix = history_set(tmp_path, cfg, 4, 4)
    rg = read_grid(prefix + "0001.nc")
    cfg2 = synth.Config(cfg.name, [rg], cfg.hydro, cfg.params, cfg.particles, cfg.refine)
    hf = HydroFiles([prefix], filenum=1, multifile=True)
    a = O.run_oracle(cfg)
    b = O.run_oracle(cfg2, records=hf)
    hf.close()
    for k in ("x", "y", "z", "status", "age"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_fortran_binding_reads_a_record(tmp_path):
    """The ISO_C_BINDING interfaces of rp_ncin_* (romspath_hip.f90) used from
    Fortran as updateHydro would: record 2 of zeta into a (xi_rho, eta_rho)
    array, checked against scipy's reading of the same file."""
    import subprocess
    from romspath_amd.build import FC, PKG
    cfg = synth.config_mini(n_particles=10)
    prefix = history_set(tmp_path, cfg, 3, 1)
    g = cfg.grids[0]
    src = tmp_path / "rd.f90"
    src.write_text(f"""program rd
  use iso_c_binding
  use romspath_hip
  implicit none
  type(c_ptr) :: f
  integer(c_int64_t) :: nt
  double precision :: z({g.L}, {g.M})
  integer :: rc
  rc = rp_ncin_open(f, '{prefix}0001.nc' // c_null_char)
  if (rc /= 0) stop 1
  rc = rp_ncin_dim(f, 'ocean_time' // c_null_char, nt)
  rc = rp_ncin_get(f, 'zeta' // c_null_char, 1_c_int64_t, z, int({g.L * g.M}, c_int64_t))
  if (rc /= 0) stop 2
  write (*, '(I4, 3ES25.17)') nt, z(1, 1), z({g.L}, 1), z(3, {g.M})
  rc = rp_ncin_close(f)
end program rd
""")
    exe = tmp_path / "rd"
    lib = PKG / "lib"
    r = subprocess.run([FC, "-O1", "-I", str(lib / "fmod"), str(src), str(lib / "fmod" / "romspath_hip.o"),
                        "-L", str(lib), "-lrp", f"-Wl,-rpath,{lib}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = r.stdout.split()
    ref = netcdf_file(prefix + "0001.nc", "r", mmap=False)
    zeta = ref.variables["zeta"][1].copy()          # (eta_rho, xi_rho)
    ref.close()
    assert int(vals[0]) == 3
    assert [float(v) for v in vals[1:]] == [zeta[0, 0], zeta[0, g.L - 1], zeta[g.M - 1, 2]]


def test_read_flags_constants_and_offsets(tmp_path):
    """updateHydro's read*/const* switches, TempOffset, and the SaltTempOn /
    WriteBottom fields (hydrodynamic_module.f90:232-512)."""
    from romspath_amd.roms import ReadFlags
    cfg = synth.config_mini(n_particles=10)
    g = cfg.grids[0]
    recs = RecordSource(cfg)
    rng = np.random.default_rng(5)
    rows = []
    for r in (1, 2):
        rec = dict(recs.get(1, r))
        rec["salt"] = 34.0 + rng.standard_normal((g.N, g.M, g.L))
        rec["temp"] = 12.0 + rng.standard_normal((g.N, g.M, g.L))
        rec["bustr"] = rng.standard_normal((g.M, g.L - 1))
        rec["bvstr"] = rng.standard_normal((g.M - 1, g.L))
        rows.append(rec)
    write_history(tmp_path / "his.nc", g, rows, [0.0, 3600.0])
    fl = ReadFlags(readZeta=False, constZeta=0.25, readW=False, constW=-1e-4, readSalt=True,
                   readTemp=True, TempOffset=1.5, SaltTempOn=True, WriteBottom=True)
    hf = HydroFiles([str(tmp_path / "his.nc")], multifile=False, flags=fl)
    got = hf.get(1, 2)
    assert np.all(got["zeta"] == 0.25) and got["zeta"].shape == (g.M, g.L)
    assert np.all(got["w"] == -1e-4) and got["w"].shape == (g.N + 1, g.M, g.L)
    assert np.array_equal(got["u"], rows[1]["u"]) and np.array_equal(got["aks"], rows[1]["aks"])
    assert np.array_equal(got["salt"], rows[1]["salt"])
    assert np.array_equal(got["temp"], rows[1]["temp"] + 1.5)
    assert np.array_equal(got["bustr"], rows[1]["bustr"]) and np.array_equal(got["bvstr"], rows[1]["bvstr"])
    fl2 = ReadFlags(SaltTempOn=True, readSalt=False, constSalt=35.0, readTemp=False, constTemp=10.0)
    got = HydroFiles([str(tmp_path / "his.nc")], multifile=False, flags=fl2).get(1, 1)
    assert np.all(got["salt"] == 35.0) and np.all(got["temp"] == 10.0)
    with pytest.raises(IndexError):
        hf.get(1, 3)
    hf.close()


def nested_history_set(tmp_path, cfg, nrec=6, recs=None, units=None, t0_days=None):
    """One history series and one Stokes series (-DSTOKES) per nest of cfg,
    records 1 .. nrec in file 0001; returns (prefixes, stokes prefixes).
    recs: the record source (default cfg's models); t0_days: write the time
    variable in days since the epoch of `units`, starting at t0_days."""
    recs = recs or RecordSource(cfg)
    prefixes, sprefixes = [], []
    for gi, g in enumerate(cfg.grids):
        d = tmp_path / f"g{gi}"
        d.mkdir()
        ids = range(1, nrec + 1)
        rows = [recs.get(gi + 1, r) for r in ids]
        if t0_days is None:
            write_history(d / "his_0001.nc", g, rows, [(r - 1) * cfg.params.dt for r in ids])
        else:
            write_history(d / "his_0001.nc", g, rows,
                          [t0_days + (r - 1) * cfg.params.dt / 86400.0 for r in ids], units=units)
        prefixes.append(str(d / "his_"))
        f = netcdf_file(str(d / "stk_0001.nc"), "w", version=2)
        f.createDimension("ocean_time", None)
        for name, n in (("xi_u", g.L - 1), ("eta_u", g.M), ("xi_v", g.L), ("eta_v", g.M - 1),
                        ("s_rho", g.N)):
            f.createDimension(name, n)
        su = f.createVariable("ustokes", "d", ("ocean_time", "s_rho", "eta_u", "xi_u"))
        sv = f.createVariable("vstokes", "d", ("ocean_time", "s_rho", "eta_v", "xi_v"))
        for r, row in enumerate(rows):
            su[r] = row["su"]
            sv[r] = row["sv"]
        f.close()
        sprefixes.append(str(d / "stk_"))
    return prefixes, sprefixes


def test_nested_oracle_run_from_files_equals_in_memory(tmp_path):
    """Two nests, one history series each (child first, as the reference
    orders them) and a Stokes series each: grids read back give the same
    scl/off, and the oracle run from the files equals the in-memory run bit
    for bit."""
    from romspath_amd.nesting import scale_offsets
    cfg = synth.config_mini_nested(n_particles=300)
    prefixes, sprefixes = nested_history_set(tmp_path, cfg)
    rgs = [read_grid(p + "0001.nc") for p in prefixes]
    s1, o1 = scale_offsets(cfg.grids, cfg.refine)
    s2, o2 = scale_offsets(rgs, cfg.refine)
    assert np.array_equal(np.asarray(s1), np.asarray(s2)) and np.array_equal(np.asarray(o1), np.asarray(o2))
    cfg2 = synth.Config(cfg.name, rgs, cfg.hydro, cfg.params, cfg.particles, cfg.refine)
    hf = HydroFiles(prefixes, filenum=1, multifile=True, stokes_prefixes=sprefixes)
    a = O.run_oracle(cfg)
    b = O.run_oracle(cfg2, records=hf)
    hf.close()
    assert set(np.unique(a.gid)) == {1, 2}
    for k in ("x", "y", "z", "status", "gid"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_wetdry_masks_from_files_drive_the_oracle_as_in_memory(tmp_path):
    """-DWETDRY: the wet/dry masks of every record read back from a history
    file (hydrodynamic_module.f90:510-570) equal what was written, and the
    oracle run from the file equals the in-memory run with the same masks,
    bit for bit."""
    cfg = synth.config_mini(n_particles=300)
    g = cfg.grids[0]
    recs = RecordSource(cfg)
    static = np.asarray(g.mask_rho, dtype=np.int32)
    nrec = 4

    def masks(r):
        wet = static.copy()
        wet[:, : 3 + r] = 0 if r % 2 else wet[:, : 3 + r]
        return dict(wetdry_mask_rho=wet, wetdry_mask_u=wet[:, :-1] * wet[:, 1:],
                    wetdry_mask_v=wet[:-1, :] * wet[1:, :])

    rows = []
    for r in range(1, nrec + 1):
        rec = dict(recs.get(1, r))
        rec.update(masks(r))
        rows.append(rec)
    write_history(tmp_path / "his_0001.nc", g, rows, [(r - 1) * cfg.params.dt for r in range(1, nrec + 1)])
    hf = HydroFiles([str(tmp_path / "his_")], filenum=1, multifile=True, wetdry=True)
    for r in range(1, nrec + 1):
        got = hf.get(1, r)
        for k, v in masks(r).items():
            assert np.array_equal(got[k], v), (r, k)

    class Mem:
        def get(self, ng, record, filenum=None):
            rec = dict(recs.get(ng, record))
            rec.update(masks(record))
            return rec

    steps = int(3 * cfg.params.dt / cfg.params.idt)
    a = O.run_oracle(cfg, max_steps=steps, records=Mem())
    b = O.run_oracle(cfg, max_steps=steps, records=hf)
    hf.close()
    for k in ("x", "y", "z", "status"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
