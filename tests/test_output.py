"""Trajectory file writer (include/rp_ncout.h; createNetCDF / writeNetCDF,
ROMSPath.f90:1140-2157) read back with scipy's independent netCDF reader,
and the getInterp2D restatement used to check the lon/lat sampler.  No GPU.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from oracle.oracle import interp2d_rho
from romspath_amd import abi
from romspath_amd.output import TrajectoryFile, meta_of
from romspath_amd.synth import RunParams


def _cols(rng, n, ngrid):
    p = {k: rng.standard_normal(n) * 100
         for k in ("age", "z", "status", "wd", "zeta", "bath", "salt", "temp", "hob", "bustr", "bvstr")}
    p["status"] = rng.choice([9.0, 0.0, 1.0, 1.1, 1.2, -8.0, -9.0], n)
    p["gid"] = rng.integers(1, ngrid + 1, n).astype(np.int32)
    return p, rng.uniform(-180, 180, n), rng.uniform(-90, 90, n)


@pytest.mark.parametrize("ngrid,extra", [(1, ()),
                                         (2, ("WaterDepth", "zeta", "h", "salttemp", "bottom",
                                              "collisions"))])
def test_records_read_back(tmp_path, ngrid, extra):
    rng = np.random.default_rng(ngrid)
    n = 1001
    prm = RunParams(idt=120, days=2.0, VTurbOn=True, deltat=12.0, seed=5, swimfast=0.0005,
                    SaltTempMean=ngrid > 1, TempOffset=1.5)
    dob = rng.uniform(0, 86400, n)
    path = tmp_path / "traj.nc"
    meta = meta_of(prm, "seconds since 2020-01-01", run_name="gyre", executable="/opt/rp",
                   institution="Rutgers")
    nc = TrajectoryFile(path, dob, ngrid, meta, extra)
    recs = []
    for r in range(4):
        p, lon, lat = _cols(rng, n, ngrid)
        nc.write(3600.0 * (r + 1), p, lon, lat)
        recs.append((p, lon, lat))
    assert nc.records == 4
    nc.close()
    f = netcdf_file(str(path), "r", mmap=False)
    try:
        assert f.version_byte == 2
        assert f.dimensions == {"numpar": n, "time": None}
        names = ["model_time", "dob", "age", "lon", "lat", "zp", "status"]
        names += ["GID"] if ngrid > 1 else []
        names += ["hitBottom", "hitLand"] if "collisions" in extra else []
        names += ["salinity", "temperature"] if "salttemp" in extra else []
        names += ["HOB", "bustr", "bvstr"] if "bottom" in extra else []
        names += [e for e in extra if e not in ("salttemp", "bottom", "collisions")]
        assert list(f.variables) == names
        v = f.variables
        assert v["model_time"].dimensions == ("time",)
        assert v["dob"].dimensions == ("numpar",)
        for k in names[2:]:
            assert v[k].dimensions == ("time", "numpar"), k
        assert v["age"].data.dtype == np.dtype(">f8") and v["lon"].data.dtype == np.dtype(">f4")
        assert v["model_time"].units == b"seconds since 2020-01-01"
        assert v["zp"].units == b"meters"
        assert np.array_equal(v["model_time"][:], 3600.0 * np.arange(1, 5))
        assert np.array_equal(v["dob"][:], dob.astype(np.float32))
        for r, (p, lon, lat) in enumerate(recs):
            assert np.array_equal(v["age"][r], p["age"])                   # double stays double
            assert np.array_equal(v["lon"][r], lon.astype(np.float32))    # round to nearest
            assert np.array_equal(v["lat"][r], lat.astype(np.float32))
            assert np.array_equal(v["zp"][r], p["z"].astype(np.float32))
            assert np.array_equal(v["status"][r], p["status"].astype(np.float32))
            if ngrid > 1:
                assert np.array_equal(v["GID"][r], p["gid"].astype(np.float32))
                assert np.array_equal(v["WaterDepth"][r], p["wd"].astype(np.float32))
                assert np.array_equal(v["zeta"][r], p["zeta"].astype(np.float32))
                assert np.array_equal(v["h"][r], p["bath"].astype(np.float32))
                assert np.array_equal(v["salinity"][r], p["salt"].astype(np.float32))
                assert np.array_equal(v["temperature"][r], p["temp"].astype(np.float32))
                for k in ("hob", "bustr", "bvstr"):
                    assert np.array_equal(v["HOB" if k == "hob" else k][r], p[k].astype(np.float32))
            if "collisions" in extra:
                # the reference's counters are never incremented
                assert not v["hitBottom"][r].any() and not v["hitLand"][r].any()
        if "collisions" in extra:
            assert v["hitBottom"].long_name == b"# of times Particle Collided with Bottom"
            assert v["hitLand"].units == b"Number of Collisions"
        a = f._attributes
        # createNetCDF's global attributes in its order (ROMSPath.f90:1681-1834)
        assert list(a)[:21] == [
            "title", "author", "adv_time_step", "days_run", "Behavior", "OpenOceanBoundary",
            "swimfast", "Vertical_Turb", "seed", "turb_time_step", "spline_error_cutoff",
            "spline_smooth_param", "subset_scale", "AKs_bakcground", "Horiz_Turb", "run_name",
            "executable", "output_loc", "administrator", "institution", "Start_datetime"]
        assert list(a)[21:23] == ["SaltTempMean", "TempOffset"]
        assert a["author"] == b"Elias Hunter (hunter@marine.rutgers.edu)"
        assert a["adv_time_step"] == 120 and a["Vertical_Turb"] == b"True"
        assert a["seed"] == 5 and a["turb_time_step"] == 12.0
        assert a["OpenOceanBoundary"] == b"TRUE" and a["Horiz_Turb"] == b"False"
        assert a["swimfast"] == 0.0005 and a["TempOffset"] == 1.5
        assert a["SaltTempMean"] == (b"TRUE" if ngrid > 1 else b"FALSE")
        assert a["run_name"] == b"gyre" and a["executable"] == b"/opt/rp"
        assert a["institution"] == b"Rutgers" and a["output_loc"] == b""
        import re
        assert re.fullmatch(rb"\d{4}/\d\d/\d\d \d\d:\d\d:\d\d", a["Start_datetime"])
        if "bottom" in extra:
            assert v["HOB"].units == b"m" and v["bustr"].units == b"newton meter-2"
            assert v["bvstr"].long_name == b"V bottom stress at minimum HOB"
    finally:
        f.close()


def test_missing_column_and_bad_path_are_errors(tmp_path):
    n = 10
    prm = RunParams()
    nc = TrajectoryFile(tmp_path / "a.nc", np.zeros(n), 2, meta_of(prm), ("zeta",))
    p, lon, lat = _cols(np.random.default_rng(0), n, 2)
    p["zeta"] = None
    with pytest.raises(abi.RpError) as e:
        nc.write(0.0, p, lon, lat)
    assert e.value.code == -1
    nc.close()
    with pytest.raises(abi.RpError) as e:
        TrajectoryFile(tmp_path / "no_such_dir" / "b.nc", np.zeros(n), 1, meta_of(prm))
    assert e.value.code == -7


def test_interp2d_restatement_pinned():
    # bilinear weights reproduce an affine field exactly (up to rounding)
    L, M = 40, 30
    i = np.arange(1, L + 1, dtype=np.float64)
    j = np.arange(1, M + 1, dtype=np.float64)
    lon = np.broadcast_to(-70.0 + 0.01 * (i - 1), (M, L)).copy()
    lat = np.broadcast_to(38.0 + 0.02 * (j - 1)[:, None], (M, L)).copy()
    rng = np.random.default_rng(1)
    X, Y = rng.uniform(1.0, L, 500), rng.uniform(1.0, M, 500)
    assert np.abs(interp2d_rho(lon, X, Y) - (-70.0 + 0.01 * (X - 1))).max() < 1e-12
    assert np.abs(interp2d_rho(lat, X, Y) - (38.0 + 0.02 * (Y - 1))).max() < 1e-12
    # a hand case on a non-affine field: corners (I,J)=1, (I+1,J)=2, (I+1,J+1)=4, (I,J+1)=8
    f = np.zeros((M, L))
    f[2, 4], f[2, 5], f[3, 5], f[3, 4] = 1.0, 2.0, 4.0, 8.0
    x, y = 0.25, 0.5
    want = (1 - x) * (1 - y) * 1 + x * (1 - y) * 2 + x * y * 4 + (1 - x) * y * 8
    assert interp2d_rho(f, 5 + x, 3 + y) == want
