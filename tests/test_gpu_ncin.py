"""A GPU run whose grid and hydro records come from ROMS history files
(netCDF classic, several files rolled over as run_External_Timestep does)
equals the run on the same grid and records in memory, bit for bit, in both
the FP64 and the FP32-field forms."""
import numpy as np
import pytest

from romspath_amd import synth
from romspath_amd.roms import HydroFiles, read_grid
from romspath_amd.runner import GpuRun

from test_ncin import history_set, nested_history_set

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_run_from_history_files_equals_in_memory(tmp_path, precision):
    cfg = synth.config_mini(n_particles=600, days=0.5)
    prefix = history_set(tmp_path, cfg, 4, 4)
    rg = read_grid(prefix + "0001.nc")
    cfg2 = synth.Config(cfg.name, [rg], cfg.hydro, cfg.params, cfg.particles, cfg.refine)
    print(f"case history files mini {precision}", flush=True)
    a = GpuRun(cfg, precision=precision)
    pa = a.run()
    a.close()
    hf = HydroFiles([prefix], filenum=1, multifile=True)
    b = GpuRun(cfg2, precision=precision, records=hf)
    pb = b.run()
    b.close()
    hf.close()
    for k in ("x", "y", "z", "status", "age", "gid"):
        assert np.array_equal(pa[k], pb[k]), k
    assert (pa["age"] > 0).sum() > 300


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_nested_run_from_history_files_equals_in_memory(tmp_path, precision):
    cfg = synth.config_mini_nested(n_particles=600)
    prefixes, sprefixes = nested_history_set(tmp_path, cfg)
    rgs = [read_grid(p + "0001.nc") for p in prefixes]
    cfg2 = synth.Config(cfg.name, rgs, cfg.hydro, cfg.params, cfg.particles, cfg.refine)
    print(f"case nested history files {precision}", flush=True)
    a = GpuRun(cfg, precision=precision)
    pa = a.run()
    a.close()
    hf = HydroFiles(prefixes, filenum=1, multifile=True, stokes_prefixes=sprefixes)
    b = GpuRun(cfg2, precision=precision, records=hf)
    pb = b.run()
    b.close()
    hf.close()
    assert set(np.unique(pa["gid"])) == {1, 2}
    for k in ("x", "y", "z", "status", "age", "gid"):
        assert np.array_equal(pa[k], pb[k]), k
