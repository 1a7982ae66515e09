"""Trajectory output on the GPU (dataOutput, ROMSPath.f90:1012-1072):
rp_sample_lonlat equals the getInterp2D restatement bit for bit, and a run's
netCDF file holds, record by record, the columns rp_get_particles returns.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from oracle.oracle import interp2d_rho
from romspath_amd import synth
from romspath_amd.output import TrajectoryFile, meta_of
from romspath_amd.runner import GpuRun

pytestmark = pytest.mark.gpu


def _want_lonlat(cfg, p):
    lon = np.full(len(p["x"]), np.nan)
    lat = np.full(len(p["x"]), np.nan)
    for gi, g in enumerate(cfg.grids):
        s = p["gid"] == gi + 1
        lon[s] = interp2d_rho(g.x_rho, p["x"][s], p["y"][s])
        lat[s] = interp2d_rho(g.y_rho, p["x"][s], p["y"][s])
    return lon, lat


def test_lonlat_sampler_matches_restatement():
    cfg = synth.config_mini_nested(n_particles=600)
    print("case nested mini fp64 40 steps, lon/lat sampler", flush=True)
    run = GpuRun(cfg, precision="fp64", sort_interval=5)
    p = run.run(max_steps=40)
    lon, lat = run.ctx.sample_lonlat()
    run.close()
    assert set(np.unique(p["gid"])) == {1, 2}
    want_lon, want_lat = _want_lonlat(cfg, p)
    assert np.array_equal(lon, want_lon)
    assert np.array_equal(lat, want_lat)


def test_run_writes_records_of_particle_columns(tmp_path):
    cfg = synth.config_mini_nested(n_particles=600)     # 2 h: records at 0, 1 h and 2 h
    print("case nested mini fp32 60 steps with trajectory file", flush=True)
    run = GpuRun(cfg, precision="fp32", sort_interval=5)
    path = tmp_path / "traj.nc"
    nc = TrajectoryFile(path, cfg.particles.dob, cfg.ngrid, meta_of(cfg.params),
                        ("WaterDepth", "zeta", "h"))
    seen = []

    def on_step(st, r):
        if st.output:
            p = r.ctx.get_particles()
            seen.append((st.ix[2], p, *r.ctx.sample_lonlat()))

    run.run(max_steps=60, output=nc, on_step=on_step)
    nc.close()
    run.close()
    assert len(seen) == 2
    f = netcdf_file(str(path), "r", mmap=False)
    try:
        v = f.variables
        # the release record at time 0 (writeNetCDF(0, ...), ROMSPath.f90:453),
        # then one per iprint at int(ix(3)); reftime is 0 for these grids
        assert np.array_equal(v["model_time"][:], [0.0] + [float(int(s[0])) for s in seen])
        assert np.array_equal(v["dob"][:], cfg.particles.dob.astype(np.float32))
        rel = cfg.particles
        assert np.array_equal(v["age"][0], rel.age)
        assert np.array_equal(v["zp"][0], rel.z.astype(np.float32))
        assert np.array_equal(v["status"][0], rel.status.astype(np.float32))
        lon0, lat0 = _want_lonlat(cfg, {"x": rel.x, "y": rel.y, "gid": rel.gid})
        assert np.array_equal(v["lon"][0], lon0.astype(np.float32))
        assert np.all(v["h"][0] == np.float32(9999.0))     # pBath before any step
        for r, (_, p, lon, lat) in enumerate(seen, start=1):
            assert np.array_equal(v["age"][r], p["age"])
            assert np.array_equal(v["zp"][r], p["z"].astype(np.float32))
            assert np.array_equal(v["status"][r], p["status"].astype(np.float32))
            assert np.array_equal(v["GID"][r], p["gid"].astype(np.float32))
            assert np.array_equal(v["lon"][r], lon.astype(np.float32))
            assert np.array_equal(v["lat"][r], lat.astype(np.float32))
            assert np.array_equal(v["WaterDepth"][r], p["wd"].astype(np.float32))
            assert np.array_equal(v["zeta"][r], p["zeta"].astype(np.float32))
            assert np.array_equal(v["h"][r], p["bath"].astype(np.float32))
            want_lon, _ = _want_lonlat(cfg, p)
            assert np.array_equal(lon, want_lon)
    finally:
        f.close()


def test_release_record_lonlat_and_reftime(tmp_path):
    """The release record writes the lon/lat it is given (the release file's
    values, writeNetCDF(0, pLon, pLat)), and every model_time is time +
    reftime of the grid (ROMSPath.f90:1931)."""
    cfg = synth.config_mini(n_particles=200, days=2 / 24.0)
    cfg.grids[-1].reftime = 86400.0 * 3
    print("case mini fp64 release record with reftime", flush=True)
    run = GpuRun(cfg, precision="fp64")
    nc = TrajectoryFile(tmp_path / "r.nc", cfg.particles.dob, cfg.ngrid, meta_of(cfg.params))
    rng = np.random.default_rng(5)
    lon = rng.uniform(-70, -60, cfg.particles.n)
    lat = rng.uniform(30, 40, cfg.particles.n)
    try:
        run.run(output=nc, release_lonlat=(lon, lat))
    finally:
        nc.close()
        run.close()
        del cfg.grids[-1].reftime
    f = netcdf_file(str(tmp_path / "r.nc"), "r", mmap=False)
    try:
        v = f.variables
        assert np.array_equal(v["model_time"][:], 86400.0 * 3 + np.array([0.0, 3600.0, 7200.0]))
        assert np.array_equal(v["lon"][0], lon.astype(np.float32))
        assert np.array_equal(v["lat"][0], lat.astype(np.float32))
    finally:
        f.close()


def test_run_writes_sampler_variables(tmp_path):
    """GpuRun.run with a trajectory file that has the SaltTempOn (mean) and
    WriteBottom variables: the runner loads salt/temp and bustr/bvstr with
    each record, starts the samplers after the first loads, updates them
    after every step, and every record holds the oracle's values (its
    restatements on the oracle's own run with the same injected deviates),
    in FP64, from the release record at time 0 on."""
    import types
    from oracle import oracle as O
    from romspath_amd.output import TrajectoryFile, meta_of
    from romspath_amd.runner import GpuRun, RecordSource
    from scipy.io import netcdf_file
    from test_bottom import stress_fields

    cfg = synth.config_mini_nested(n_particles=400)

    class Recs(RecordSource):
        def get(self, ng, record, filenum=None):
            rec = dict(super().get(ng, record, filenum))
            g = cfg.grids[ng - 1]
            rng = np.random.default_rng(100 * ng + record)
            rec["salt"] = 34.0 + rng.standard_normal((g.N, g.M, g.L))
            rec["temp"] = 12.0 + rng.standard_normal((g.N, g.M, g.L))
            rec["bustr"], rec["bvstr"] = stress_fields(g, 3600.0 * record, rng)
            return rec

    recs = Recs(cfg)
    print("case run with sampler variables, mini-nested fp64", flush=True)
    run = GpuRun(cfg, precision="fp64", records=recs, sort_interval=4)
    nc = TrajectoryFile(tmp_path / "t.nc", cfg.particles.dob, cfg.ngrid, meta_of(cfg.params),
                        ("salttemp", "bottom"))
    state = {}
    snaps = []

    def snapshot():
        snaps.append((state["ms"] / state["mI"], state["mt"] / state["mI"], state["hob"].copy(),
                      state["bu"].copy(), state["bv"].copy()))
        state["ms"][:] = 0.0
        state["mt"][:] = 0.0
        state["mI"] = 0
        state["hob"][:] = 9999.0

    def before_first(st, s, grids):
        # ini_ROMSPath: the release samples, then the release record
        state["grids"] = grids
        la, ps, pt, ms, mt, mI = O.tracers_begin(grids, st.tc, s, True)
        state.update(la=la, ps=ps, pt=pt, ms=ms, mt=mt, mI=mI, bla=s.age.copy(),
                     hob=np.full(s.n, 9999.0), bu=np.zeros(s.n), bv=np.zeros(s.n))
        snapshot()

    def on_step(st, s):
        grids = state["grids"]
        T = O.make_time(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf)
        O.tracers_update(grids, st, s, state["la"], state["ps"], state["pt"], state["ms"], state["mt"])
        state["mI"] += 1
        O.bottom_update(grids, T, s, state["bla"], state["hob"], state["bu"], state["bv"])
        if st.output:
            snapshot()

    try:
        run.run(output=nc, inject=True, salttemp_mean=True)
    finally:
        nc.close()
        run.close()
    # the oracle's samplers on the oracle's own run with the same deviates
    O.run_oracle(cfg, inject=True, records=recs, on_step=on_step, before_first=before_first)
    f = netcdf_file(str(tmp_path / "t.nc"), "r", mmap=False)
    try:
        v = f.variables
        assert len(snaps) == v["model_time"].shape[0] >= 2
        for r, (sa, te, hob, bu, bv) in enumerate(snaps):
            assert np.array_equal(v["salinity"][r], sa.astype(np.float32))
            assert np.array_equal(v["temperature"][r], te.astype(np.float32))
            assert np.array_equal(v["HOB"][r], hob.astype(np.float32))
            assert np.array_equal(v["bustr"][r], bu.astype(np.float32))
            assert np.array_equal(v["bvstr"][r], bv.astype(np.float32))
        assert (v["HOB"][-1] < 9999.0).sum() > 200
    finally:
        f.close()
