"""The Fortran host (romspath_hip module + rp_driver) drives librp.so through
ISO_C_BINDING and reproduces the oracle on the CPU-runnable double gyre."""
import subprocess

import numpy as np
import pytest

from scipy.io import netcdf_file

from oracle import oracle as O
from oracle.oracle import interp2d_rho
from romspath_amd import synth
from romspath_amd.build import DRIVER

NML = """&numparticles numpar={n} /
&timeparam days={days}, iprint=3600, dt=3600, idt=60 /
&turbparam VTurbOn=.FALSE., serr=1.0D-4, smth=0.6D0, sub=4.0D0, deltat=1.0D0,
  AKSback=1.0D-8, HTurbOn=.FALSE., ConstantHTurb=0.0D0 /
&advectparam scheme=1, nsb=0, vertdist=0.0D0 /
&romsgrid Ngrid=1, refine=1.0D0 /
"""


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    n, days = 500, 3 / 24.0
    cfg = synth.config1(n_particles=n, days=days)
    p = cfg.particles
    (tmp_path / "run.nml").write_text(NML.format(n=n, days=days))
    # the reference's release file: lon lat z dob, here x_rho / y_rho in metres
    g = cfg.grids[0]
    lon, lat = (p.x - 1.0) * g.dx, (p.y - 1.0) * g.dy
    np.savetxt(tmp_path / "par.txt", np.c_[lon, lat, p.z, p.dob], fmt="%.17e")
    print("case fortran driver C1", n, "particles", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path / "run.nml"), str(tmp_path / "par.txt"),
                        str(tmp_path / "out.txt"), str(tmp_path / "traj.nc")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.loadtxt(tmp_path / "out.txt")
    # the oracle from the oracle's own placement of the same release file
    loc = O.locate(cfg, lon, lat, p.z)
    assert loc["isin"].all() and np.abs(loc["x"] - p.x).max() < 1e-9
    placed = synth.Particles(loc["x"], loc["y"], loc["z"], p.dob, gid=loc["gid"],
                             status=loc["status"], isin=loc["isin"])
    o = O.run_oracle(cfg, particles=placed)
    assert np.array_equal(got[:, 3], o.status)
    assert np.abs(got[:, 0] - o.x).max() <= 1e-9
    assert np.abs(got[:, 1] - o.y).max() <= 1e-9
    assert np.abs(got[:, 2] - o.z).max() <= 1e-9 * 50.0 / 20
    # the trajectory file: one record per hour (iprint = 3600 s), the last
    # one holding the final state; lon/lat are getInterp2D of x_rho/y_rho
    f = netcdf_file(str(tmp_path / "traj.nc"), "r", mmap=False)
    try:
        v = f.variables
        assert np.array_equal(v["model_time"][:], [3600.0, 7200.0, 10800.0])
        assert np.array_equal(v["dob"][:], p.dob.astype(np.float32))
        assert np.array_equal(v["zp"][2], got[:, 2].astype(np.float32))
        assert np.array_equal(v["status"][2], got[:, 3].astype(np.float32))
        assert np.array_equal(v["lon"][2], interp2d_rho(g.x_rho, got[:, 0], got[:, 1]).astype(np.float32))
        assert np.array_equal(v["lat"][2], interp2d_rho(g.y_rho, got[:, 0], got[:, 1]).astype(np.float32))
        assert "GID" not in v
    finally:
        f.close()


def test_fortran_driver_is_built():
    assert DRIVER.exists()


@pytest.mark.gpu
def test_fortran_driver_reads_history_files(tmp_path):
    """The driver's $romsoutput mode: grid and records from ROMS history files
    (netCDF classic, 4 records a file, rolled over to the second file at the
    fifth external step) through rp_ncin; the result equals the oracle's on
    the same records."""
    from test_ncin import history_set
    n, days = 300, 6 / 24.0
    cfg = synth.config1(n_particles=n, days=days)
    p = cfg.particles
    prefix = history_set(tmp_path, cfg, 4, 2)
    nml = NML.format(n=n, days=days) + (
        f"&romsoutput prefix='{prefix}', suffix='.nc', filenum=1, numdigits=4, multifile=.TRUE.,\n"
        "  time_vname='ocean_time', time_dname='ocean_time' /\n")
    (tmp_path / "run.nml").write_text(nml)
    g = cfg.grids[0]
    lon, lat = (p.x - 1.0) * g.dx, (p.y - 1.0) * g.dy
    np.savetxt(tmp_path / "par.txt", np.c_[lon, lat, p.z, p.dob], fmt="%.17e")
    print("case fortran driver from history files", n, "particles", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path / "run.nml"), str(tmp_path / "par.txt"),
                        str(tmp_path / "out.txt")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.loadtxt(tmp_path / "out.txt")
    loc = O.locate(cfg, lon, lat, p.z)
    placed = synth.Particles(loc["x"], loc["y"], loc["z"], p.dob, gid=loc["gid"],
                             status=loc["status"], isin=loc["isin"])
    o = O.run_oracle(cfg, particles=placed)
    assert np.array_equal(got[:, 3], o.status)
    assert np.abs(got[:, 0] - o.x).max() <= 1e-9
    assert np.abs(got[:, 1] - o.y).max() <= 1e-9
    assert np.abs(got[:, 2] - o.z).max() <= 1e-9 * 50.0 / 20
