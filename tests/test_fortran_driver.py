"""The Fortran host (romspath_hip module + rp_driver) drives librp.so through
ISO_C_BINDING as the reference's main program does: it reads a namelist file
with all sixteen groups of getParams (parameter_module.f90:41-104), the
reference's own ROMSPath.data among them, and runs from it.  On the GPU it
reproduces the oracle on the CPU-runnable double gyre and on one grid read
from history files, and on nested grids from files with Stokes drift, both
samplers and every optional output variable it writes the same file as the
Python harness driving the same library."""
import dataclasses
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy.io import netcdf_file

from oracle import oracle as O
from oracle.oracle import interp2d_rho
from romspath_amd import synth
from romspath_amd.build import DRIVER

REF_DATA = Path("/root/reference/ROMSPath.data")


def _f(v):
    if isinstance(v, bool):
        return ".TRUE." if v else ".FALSE."
    if isinstance(v, str):
        return f"'{v}'"
    if isinstance(v, (list, tuple)):
        return ", ".join(_f(x) for x in v)
    if isinstance(v, float):
        return repr(v).replace("e", "D") if "e" in repr(v) else repr(v) + "D0"
    return str(v)


def namelist(**over):
    """Every group of getParams, in its order, with every variable of
    ROMSPath.h:69-345; `over` replaces values by name."""
    groups = [
        ("numparticles", dict(numpar=100)),
        ("timeparam", dict(days=0.125, iprint=3600, dt=3600, idt=60)),
        ("hydroparam", dict(readZeta=True, constZeta=0.0, readSalt=True, constSalt=0.0,
                            readTemp=True, constTemp=0.0, readU=True, constU=0.0, readV=True,
                            constV=0.0, readW=True, constW=0.0, readAks=True, constAks=0.0,
                            readDens=False, constDens=0.0, stokesprefix=["", ""],
                            turbstd_v_a_prefix=["", ""], wavestd_prefix=["", ""],
                            Process_VA=False, Process_WA=False)),
        ("turbparam", dict(VTurbOn=False, serr=1.0e-4, smth=0.6, sub=4.0, deltat=1.0,
                           AKSback=1.0e-8, HTurbOn=False, ConstantHTurb=[0.0] * 5)),
        ("advectparam", dict(scheme=1, nsb=0, vertdist=0.0)),
        ("behavparam", dict(Behavior=0, OpenOceanBoundary=True, pediage=302400.0, swimstart=0.0,
                            swimslow=0.0005, swimfast=0.0005, Sgradient=1.0, sink=-0.0003,
                            Hswimspeed=0.9, Swimdepth=2.0)),
        ("fuchsparam", dict(vort_cr=1.0, vort_sat=10.0, b0pv=0.5, b1pv=1.0, b0wv=0.0, b1w=0.0,
                            acc_cr=10.0, acc_sat=1000.0, b0pa=0.5, b1pa=1.0, b0wa=0.2, va_flag=1)),
        ("growparam", dict(Growth=0, mortality=True, deadage=3888000.0, initsize=250.0,
                           maxsize=1000.0, tempcut=2.0, a0=-29.8, a1=3.86, a2=0.0, a3=0.0,
                           a4=-0.07, a5=0.0, a6=0.0, a7=0.0, a8=0.0)),
        ("dvmparam", dict(twistart=4.801821, twiend=19.19956, daylength=14.39774, Em=1814.328,
                          Kd=1.07, thresh=0.0166)),
        ("settleparam", dict(settlementon=False, holesExist=True, minpolyid=101001,
                             maxpolyid=101004, minholeid=100201, maxholeid=100401, pedges=67,
                             hedges=32)),
        ("romsgrid", dict(Ngrid=1, refine=[1.0])),
        ("romsoutput", dict(prefix=[""], suffix=".nc", filenum=1, numdigits=4, multifile=False,
                            time_vname="ocean_time", time_dname="ocean_time")),
        ("parloc", dict(parfile="par.txt")),
        ("HabPolyLoc", dict(habitatfile="./input/End_polygons.csv",
                            holefile="./input/End_holes.csv")),
        ("output", dict(outpath="./", NCOutFile="traj", outpathGiven=True, NCtime=0,
                        RunName="driver test", ExeDir="/opt/romspath", OutDir="./output",
                        RunBy="a tester", Institution="RUTGERS", StartedOn="today")),
        ("other", dict(seed=1, SaltTempOn=False, TrackCollisions=False, WriteHeaders=False,
                       SaltTempMean=False, WriteBottom=False, WriteModelTiming=False,
                       WriteProblemFile=False, ijbuff=4, ErrorFlag=0, FreeSlip=False,
                       TempOffset=0.0, WriteWaterDepth=False, WriteZeta=False, WriteBath=False)),
    ]
    used = set()
    lines = ["! every group of getParams, in its order"]
    for name, vals in groups:
        lines.append(f"&{name}")
        for k, v in vals.items():
            if k in over:
                v = over[k]
                used.add(k)
            lines.append(f"  {k} = {_f(v)}")
        lines.append("/")
    assert used == set(over), set(over) - used
    return "\n".join(lines) + "\n"


def _params(text):
    out = {}
    for line in text.splitlines():
        for m in re.finditer(r"(\w+(?:\(\d\))?) =[ ]+(\S+)", line):
            out[m.group(1)] = m.group(2)
    return out


def _run(args, cwd, timeout=120):
    r = subprocess.run([str(DRIVER), *map(str, args)], capture_output=True, text=True,
                       timeout=timeout, cwd=str(cwd))
    print(r.stdout[-3000:], r.stderr[-3000:])
    return r


def test_fortran_driver_is_built():
    assert DRIVER.exists()


@pytest.mark.skipif(not REF_DATA.exists(), reason="the reference's ROMSPath.data is not here")
def test_reads_reference_namelist_unchanged():
    """The reference's ROMSPath.data ($group ... $end, comments, multi-line
    arrays) reads through all sixteen groups (-n: parse and print, no GPU)."""
    r = _run(["-n", REF_DATA], "/tmp")
    assert r.returncode == 0
    p = _params(r.stdout)
    assert p["numpar"] == "3000" and p["dt"] == "720" and p["idt"] == "60"
    assert p["Ngrid"] == "2" and float(p["refine"]) == 7.0
    assert p["prefix(2)"].endswith("doppio_his_") and "prefix(3)" not in p
    assert p["stokesprefix(1)"].endswith("snaildel_stokes_")
    assert p["time_vname"] == "time" and p["filenum"] == "2" and p["multifile"] == "F"
    assert p["parfile"] == "./init_nested.csv" and p["NCOutFile"] == "SHOWNESTING_NOMIX"
    assert p["seed"] == "0" and p["ErrorFlag"] == "3"
    assert float(p["swimfast"]) == 0.0005 and float(p["thresh"]) == 0.0166
    assert "RunBy = ELI HUNTER" in r.stdout
    assert "SaltTempOn SaltTempMean WriteBottom WriteWaterDepth WriteZeta WriteBath =  T T T T F F" \
        in r.stdout


def test_reads_every_group_and_reports_missing_ones(tmp_path):
    (tmp_path / "all.nml").write_text(namelist(numpar=7, dt=720, Ngrid=2, refine=[3.0],
                                               prefix=["a_", "b_"], seed=11, WriteZeta=True))
    r = _run(["-n", "all.nml"], tmp_path)
    assert r.returncode == 0
    p = _params(r.stdout)
    assert p["numpar"] == "7" and p["dt"] == "720" and p["Ngrid"] == "2"
    assert p["prefix(1)"] == "a_" and p["prefix(2)"] == "b_" and p["seed"] == "11"
    assert float(p["refine"]) == 3.0 and p["hedges"] == "32"
    # a group left out: getParams' message for it, and a failed exit
    text = namelist()
    cut = text.index("&dvmparam")
    (tmp_path / "short.nml").write_text(text[:cut] + text[text.index("&settleparam"):])
    r = _run(["-n", "short.nml"], tmp_path)
    assert r.returncode != 0
    assert "Error when reading growthparam, pls check ROMSPath.data" in r.stdout
    # a variable the group does not have
    (tmp_path / "bad.nml").write_text(text.replace("  nsb = 0", "  nsb = 0\n  nosuch = 1"))
    r = _run(["-n", "bad.nml"], tmp_path)
    assert r.returncode != 0 and "Error when reading advectparam" in r.stdout


def _release_file(path, cfg):
    """The reference's release file: lon lat z dob, lon/lat getInterp2D of
    each particle's grid's x_rho / y_rho at its X, Y."""
    p = cfg.particles
    lon, lat = np.empty(p.n), np.empty(p.n)
    for gi, g in enumerate(cfg.grids):
        s = p.gid == gi + 1
        lon[s] = interp2d_rho(g.x_rho, p.x[s], p.y[s])
        lat[s] = interp2d_rho(g.y_rho, p.x[s], p.y[s])
    np.savetxt(path, np.c_[lon, lat, p.z, p.dob], fmt="%.17e")
    # the values the driver reads back
    a = np.loadtxt(path)
    return a[:, 0], a[:, 1], a[:, 2], a[:, 3]


def _state(path):
    a = np.loadtxt(path)
    return {"x": a[:, 0], "y": a[:, 1], "z": a[:, 2], "status": a[:, 3], "gid": a[:, 4].astype(int),
            "age": a[:, 5]}


def _oracle_from_release(cfg, lon, lat, z, dob, records=None):
    loc = O.locate(cfg, lon, lat, z, records=records)
    placed = synth.Particles(loc["x"], loc["y"], loc["z"], dob, gid=loc["gid"],
                             status=loc["status"], isin=loc["isin"])
    return loc, O.run_oracle(cfg, particles=placed, records=records)


def _check_vs_oracle(got, o):
    assert np.array_equal(got["status"], o.status)
    assert np.abs(got["x"] - o.x).max() <= 1e-9
    assert np.abs(got["y"] - o.y).max() <= 1e-9
    assert np.abs(got["z"] - o.z).max() <= 1e-9 * 50.0 / 20


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    """The analytic double gyre (prefix blank): oracle parity of the final
    state, and the trajectory file of $output with the release record at
    time 0, one record per iprint, and createNetCDF's global attributes
    taken from the namelist."""
    n, days = 500, 3 / 24.0
    cfg = synth.config1(n_particles=n, days=days)
    lon, lat, z, dob = _release_file(tmp_path / "par.txt", cfg)
    (tmp_path / "run.nml").write_text(namelist(numpar=n, days=days, parfile="par.txt",
                                               outpath=str(tmp_path) + "/", swimfast=0.0025))
    print("case fortran driver C1", n, "particles", flush=True)
    r = _run(["run.nml", "state.txt"], tmp_path)
    assert r.returncode == 0
    got = _state(tmp_path / "state.txt")
    loc, o = _oracle_from_release(cfg, lon, lat, z, dob)
    assert loc["isin"].all() and np.abs(loc["x"] - cfg.particles.x).max() < 1e-9
    _check_vs_oracle(got, o)
    g = cfg.grids[0]
    f = netcdf_file(str(tmp_path / "traj.nc"), "r", mmap=False)
    try:
        v = f.variables
        assert list(v) == ["model_time", "dob", "age", "lon", "lat", "zp", "status"]
        assert np.array_equal(v["model_time"][:], [0.0, 3600.0, 7200.0, 10800.0])
        assert np.array_equal(v["dob"][:], dob.astype(np.float32))
        # writeNetCDF(0, pLon, pLat): the release file's lon/lat
        assert np.array_equal(v["lon"][0], lon.astype(np.float32))
        assert np.array_equal(v["lat"][0], lat.astype(np.float32))
        assert np.array_equal(v["zp"][0], loc["z"].astype(np.float32))
        assert np.array_equal(v["status"][0], loc["status"].astype(np.float32))
        assert np.array_equal(v["zp"][3], got["z"].astype(np.float32))
        assert np.array_equal(v["status"][3], got["status"].astype(np.float32))
        assert np.array_equal(v["lon"][3], interp2d_rho(g.x_rho, got["x"], got["y"]).astype(np.float32))
        assert np.array_equal(v["lat"][3], interp2d_rho(g.y_rho, got["x"], got["y"]).astype(np.float32))
        a = f._attributes
        assert a["author"] == b"Elias Hunter (hunter@marine.rutgers.edu)"
        assert a["swimfast"] == 0.0025 and a["OpenOceanBoundary"] == b"TRUE"
        assert a["run_name"] == b"driver test" and a["administrator"] == b"a tester"
        assert a["executable"] == b"/opt/romspath" and a["institution"] == b"RUTGERS"
        assert a["SaltTempMean"] == b"FALSE" and a["TempOffset"] == 0.0
        assert v["model_time"].units == b"seconds since model start"
    finally:
        f.close()


@pytest.mark.gpu
def test_fortran_driver_reads_history_files(tmp_path):
    """One grid from ROMS history files (4 records a file, rolled over at the
    fifth external step; the time variable in days from day 2): the result
    equals the oracle's on the same records; model_time is time + reftime in
    the converted units; NCtime /= 0 without outpath names the file
    NCOutFile_001.nc in the working directory."""
    from test_ncin import write_history
    from romspath_amd.runner import RecordSource
    n, days = 300, 6 / 24.0
    cfg = synth.config1(n_particles=n, days=days)
    recs = RecordSource(cfg)
    g = cfg.grids[0]
    for k in range(2):
        ids = range(4 * k + 1, 4 * k + 5)
        write_history(tmp_path / f"his_{k + 1:04d}.nc", g, [recs.get(1, r) for r in ids],
                      [2.0 + (r - 1) * 3600.0 / 86400.0 for r in ids],
                      units="days since 2000-01-01 00:00:00")
    lon, lat, z, dob = _release_file(tmp_path / "par.txt", cfg)
    (tmp_path / "run.nml").write_text(namelist(
        numpar=n, days=days, parfile="par.txt", prefix=[str(tmp_path / "his_")], multifile=True,
        outpathGiven=False, NCtime=3600, NCOutFile="run"))
    print("case fortran driver from history files", n, "particles", flush=True)
    r = _run(["run.nml", "state.txt"], tmp_path)
    assert r.returncode == 0
    got = _state(tmp_path / "state.txt")
    _, o = _oracle_from_release(cfg, lon, lat, z, dob)
    _check_vs_oracle(got, o)
    f = netcdf_file(str(tmp_path / "run_001.nc"), "r", mmap=False)
    try:
        v = f.variables
        assert np.array_equal(v["model_time"][:], 2 * 86400.0 + 3600.0 * np.arange(7))
        assert v["model_time"].units == b"seconds since 2000-01-01 00:00:00"
    finally:
        f.close()


@pytest.mark.gpu
def test_fortran_driver_nested_files_samplers_equal_harness(tmp_path):
    """Two nests from history files with a Stokes series each (--stokes),
    VTurb and HTurb on the library's RNG, SaltTempOn with SaltTempMean and
    TempOffset, WriteBottom, WriteWaterDepth / WriteZeta / WriteBath and
    TrackCollisions: the driver's trajectory file equals, variable by
    variable and bit for bit, the one the Python harness writes driving the
    same library from the same files and namelist values, and the final
    states are equal."""
    from test_bottom import stress_fields
    from test_ncin import nested_history_set
    from romspath_amd.output import TrajectoryFile, meta_of
    from romspath_amd.roms import HydroFiles, ReadFlags, read_grid
    from romspath_amd.runner import GpuRun, RecordSource

    cfg = synth.config_mini_nested(n_particles=400, days=0.125)

    class Recs(RecordSource):
        def get(self, ng, record, filenum=None):
            rec = dict(super().get(ng, record, filenum))
            gg = cfg.grids[ng - 1]
            rng = np.random.default_rng(100 * ng + record)
            rec["salt"] = 34.0 + rng.standard_normal((gg.N, gg.M, gg.L))
            rec["temp"] = 12.0 + rng.standard_normal((gg.N, gg.M, gg.L))
            rec["bustr"], rec["bvstr"] = stress_fields(gg, 3600.0 * record, rng)
            return rec

    units = "days since 2000-01-01 00:00:00"
    prefixes, sprefixes = nested_history_set(tmp_path, cfg, recs=Recs(cfg), units=units, t0_days=3.5)
    lon, lat, z, dob = _release_file(tmp_path / "par.txt", cfg)
    prm = dataclasses.replace(cfg.params, SaltTempMean=True, TempOffset=0.5, swimfast=0.001)
    (tmp_path / "run.nml").write_text(namelist(
        numpar=cfg.particles.n, days=prm.days, dt=prm.dt, idt=prm.idt, VTurbOn=True, HTurbOn=True,
        deltat=prm.deltat, ConstantHTurb=list(prm.ConstantHTurb), Ngrid=2,
        refine=[float(cfg.refine[0])], prefix=prefixes, stokesprefix=sprefixes, multifile=True,
        parfile="par.txt", outpath=str(tmp_path) + "/", seed=prm.seed, SaltTempOn=True,
        SaltTempMean=True, TempOffset=0.5, WriteBottom=True, WriteWaterDepth=True,
        WriteZeta=True, WriteBath=True, TrackCollisions=True, swimfast=0.001))
    print("case fortran driver nested files with samplers", flush=True)
    r = _run(["--stokes", "run.nml", "state.txt"], tmp_path)
    assert r.returncode == 0
    got = _state(tmp_path / "state.txt")

    # the Python harness on the same files and values
    rgs = [read_grid(p + "0001.nc") for p in prefixes]
    assert rgs[-1].reftime == 3.5 * 86400.0
    cfg2 = synth.Config(cfg.name, rgs, cfg.hydro, prm, cfg.particles, cfg.refine)
    flags = ReadFlags(readSalt=True, readTemp=True, TempOffset=0.5, SaltTempOn=True, WriteBottom=True)
    hf = HydroFiles(prefixes, filenum=1, multifile=True, flags=flags, stokes_prefixes=sprefixes)
    run = GpuRun(cfg2, precision="fp64", records=hf)
    path = tmp_path / "harness.nc"
    try:
        run.load_initial()
        loc = run.ctx.locate_particles(lon, lat, z)
        placed = synth.Particles(loc["x"], loc["y"], loc["z"], dob, gid=loc["gid"],
                                 status=loc["status"], isin=loc["isin"])
        run.ctx.set_particles(placed)
        nc = TrajectoryFile(path, dob, 2, meta_of(prm, rgs[-1].time_units),
                            ("WaterDepth", "zeta", "h", "salttemp", "bottom", "collisions"))
        try:
            p = run.run(output=nc, salttemp_mean=True, release_lonlat=(lon, lat))
        finally:
            nc.close()
    finally:
        run.close()
        hf.close()
    assert set(np.unique(loc["gid"])) == {1, 2}
    for k in ("x", "y", "z", "status", "age"):
        assert np.array_equal(got[k], p[k]), k
    assert np.array_equal(got["gid"], p["gid"])
    fa = netcdf_file(str(tmp_path / "traj.nc"), "r", mmap=False)
    fb = netcdf_file(str(path), "r", mmap=False)
    try:
        va, vb = fa.variables, fb.variables
        assert list(va) == list(vb)
        assert "hitBottom" in va and "HOB" in va and "salinity" in va and "GID" in va
        for k in va:
            assert np.array_equal(va[k][:], vb[k][:]), k
        assert va["model_time"].shape[0] == 4
        assert np.array_equal(va["model_time"][:], 3.5 * 86400.0 + 3600.0 * np.arange(4))
        assert va["model_time"].units == b"seconds since 2000-01-01 00:00:00"
        a = fa._attributes
        assert a["SaltTempMean"] == b"TRUE" and a["TempOffset"] == 0.5 and a["swimfast"] == 0.001
        assert a["Vertical_Turb"] == b"True" and a["Horiz_Turb"] == b"True"
        assert a["seed"] == prm.seed
        # the samplers ran: salinity means and minimum heights above bottom
        assert (va["HOB"][-1] < 9999.0).sum() > 200
        assert abs(np.median(va["salinity"][-1][loc["isin"] == 1]) - 34.0) < 0.5
    finally:
        fa.close()
        fb.close()


@pytest.mark.gpu
def test_fortran_driver_wetdry_masks_from_files(tmp_path):
    """--wetdry: the driver reads each record's wetdry_mask_rho/_u/_v from
    the history file into its slot (the reference built with -DWETDRY,
    hydrodynamic_module.f90:510-570); the final state equals the oracle's on
    the same records and masks (a shore band that dries on odd records, on
    the island grid), and the masks change the result.  With --fp32 the
    final state is within 1e-4 cell of the same oracle run."""
    from test_ncin import write_history
    from test_gpu_wetdry import wet_masks
    from romspath_amd.roms import HydroFiles
    from romspath_amd.runner import RecordSource
    n, days = 300, 4 / 24.0
    cfg = synth.config_mini(n_particles=n, days=days, vturb=False, hturb=False, dob_spread=0.0)
    recs = RecordSource(cfg)
    g = cfg.grids[0]
    rows = []
    for r in range(1, 8):
        rec = dict(recs.get(1, r))
        rec.update(wet_masks(g, r))
        rows.append(rec)
    write_history(tmp_path / "his_0001.nc", g, rows, [(r - 1) * 3600.0 for r in range(1, 8)])
    lon, lat, z, dob = _release_file(tmp_path / "par.txt", cfg)
    p = cfg.params
    (tmp_path / "run.nml").write_text(namelist(
        numpar=n, days=days, dt=p.dt, idt=p.idt, iprint=p.iprint, parfile="par.txt",
        prefix=[str(tmp_path / "his_")], multifile=True, outpath=str(tmp_path) + "/"))
    print("case fortran driver --wetdry mini", n, "particles", flush=True)
    r = _run(["--wetdry", "run.nml", "state.txt"], tmp_path)
    assert r.returncode == 0
    got = _state(tmp_path / "state.txt")
    hf = HydroFiles([str(tmp_path / "his_")], filenum=1, multifile=True, wetdry=True)
    _, o = _oracle_from_release(cfg, lon, lat, z, dob, records=hf)
    hf.close()
    _check_vs_oracle(got, o)
    hf = HydroFiles([str(tmp_path / "his_")], filenum=1, multifile=True)
    _, plain = _oracle_from_release(cfg, lon, lat, z, dob, records=hf)
    hf.close()
    assert (np.abs(plain.x - got["x"]) + np.abs(plain.y - got["y"]) > 1e-6).sum() >= 5
    # the FP32-field form under the same masks: within 1e-4 cell of the oracle
    r = _run(["--wetdry", "--fp32", "run.nml", "state32.txt"], tmp_path)
    assert r.returncode == 0
    g32 = _state(tmp_path / "state32.txt")
    same = g32["status"] == o.status
    assert same.mean() >= 0.99
    assert np.abs(g32["x"] - o.x)[same].max() <= 1e-4 and np.abs(g32["y"] - o.y)[same].max() <= 1e-4
