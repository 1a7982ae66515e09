"""The netCDF classic reader and the trajectory writer (rp_ncin.cpp,
rp_ncout.cpp) built with AddressSanitizer and UndefinedBehaviorSanitizer
(romspath_amd/build.py build_nc_check), fed a valid history file and
hundreds of damaged copies: truncated at every part of the header and the
data, byte flips in the header, and a header whose dimension lengths and
offsets are set to extreme values.  Every call must return a code; any
out-of-bounds access, leak or undefined behaviour aborts the checker.  No GPU.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from romspath_amd import build as B

VARS = ["h", "mask_rho", "s_rho", "Cs_r", "ocean_time", "zeta", "u", "v", "temp", "nope"]


@pytest.fixture(scope="module")
def checker():
    if not B.NC_CHECK.exists() or B.NC_CHECK.stat().st_mtime < max(
            (B.CSRC / f).stat().st_mtime for f in ("rp_ncin.cpp", "rp_ncout.cpp", "rp_nc_check.cpp")):
        B.build_nc_check(verbose=False)
    return str(B.NC_CHECK)


def _history(path, version):
    L, M, N, T = 9, 7, 4, 3
    rng = np.random.default_rng(1)
    f = netcdf_file(str(path), "w", version=version)
    f.title = "damaged-input test"
    f.createDimension("ocean_time", None)
    f.createDimension("xi_rho", L)
    f.createDimension("eta_rho", M)
    f.createDimension("xi_u", L - 1)
    f.createDimension("eta_v", M - 1)
    f.createDimension("s_rho", N)
    v = f.createVariable("h", "d", ("eta_rho", "xi_rho"))
    v[:] = rng.uniform(10, 50, (M, L))
    v = f.createVariable("mask_rho", "i", ("eta_rho", "xi_rho"))
    v[:] = 1
    f.createVariable("s_rho", "d", ("s_rho",))[:] = np.linspace(-0.9, -0.1, N)
    f.createVariable("Cs_r", "f", ("s_rho",))[:] = np.linspace(-0.9, -0.1, N)
    t = f.createVariable("ocean_time", "d", ("ocean_time",))
    t.units = "seconds since 2000-01-01 00:00:00"
    z = f.createVariable("zeta", "f", ("ocean_time", "eta_rho", "xi_rho"))
    u = f.createVariable("u", "d", ("ocean_time", "s_rho", "eta_rho", "xi_u"))
    vv = f.createVariable("v", "h", ("ocean_time", "s_rho", "eta_v", "xi_rho"))
    tp = f.createVariable("temp", "b", ("ocean_time", "s_rho", "eta_rho", "xi_rho"))
    for k in range(T):
        t[k] = 3600.0 * k
        z[k] = rng.standard_normal((M, L))
        u[k] = rng.standard_normal((N, M, L - 1))
        vv[k] = rng.integers(-100, 100, (N, M - 1, L))
        tp[k] = rng.integers(-100, 100, (N, M, L))
    f.close()
    return path.read_bytes()


def _run(checker, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=87")
    return subprocess.run([checker, *args], capture_output=True, text=True, env=env, timeout=120)


@pytest.mark.parametrize("version", [1, 2])
def test_damaged_files_return_codes(tmp_path, checker, version):
    good = _history(tmp_path / "good.nc", version)
    r = _run(checker, ["read", str(tmp_path / "good.nc"), *VARS])
    assert r.returncode == 0, r.stdout + r.stderr
    variants = []
    n = len(good)
    # truncations across the header and into the data
    for cut in sorted(set([0, 1, 3, 4, 7, 8, 12, 16, 20, 31, 32, 33, 47, 64, 100]
                          + list(range(0, min(n, 1200), 13)) + [n // 2, n - 1])):
        variants.append(good[:cut])
    # byte flips and overwrites in the header
    rng = np.random.default_rng(7)
    hdr = min(n, 1100)
    for _ in range(300):
        b = bytearray(good)
        for _ in range(int(rng.integers(1, 4))):
            i = int(rng.integers(0, hdr))
            b[i] = int(rng.integers(0, 256))
        variants.append(bytes(b))
    # extreme 32-bit fields at every word boundary of the header
    for i in range(4, min(n, 700), 4):
        for val in (b"\x7f\xff\xff\xff", b"\xff\xff\xff\xff", b"\x80\x00\x00\x00"):
            b = bytearray(good)
            b[i:i + 4] = val
            variants.append(bytes(b))
    paths = []
    for k, data in enumerate(variants):
        p = tmp_path / f"bad{k}.nc"
        p.write_bytes(data)
        paths.append(p)
    for p in paths:
        r = _run(checker, ["read", str(p), *VARS])
        assert r.returncode == 0, f"{p.name}: exit {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-4000:]}"


def test_writer_round_trip_clean(tmp_path, checker):
    r = _run(checker, ["write", str(tmp_path / "traj.nc")])
    assert r.returncode == 0, r.stdout + r.stderr
