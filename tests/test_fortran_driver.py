"""The Fortran host (romspath_hip module + rp_driver) drives librp.so through
ISO_C_BINDING and reproduces the oracle on the CPU-runnable double gyre."""
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
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
    np.savetxt(tmp_path / "par.txt", np.c_[p.x, p.y, p.z, p.dob], fmt="%.17e")
    print("case fortran driver C1", n, "particles", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path / "run.nml"), str(tmp_path / "par.txt"),
                        str(tmp_path / "out.txt")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.loadtxt(tmp_path / "out.txt")
    o = O.run_oracle(cfg)
    assert np.array_equal(got[:, 3], o.status)
    assert np.abs(got[:, 0] - o.x).max() <= 1e-9
    assert np.abs(got[:, 1] - o.y).max() <= 1e-9
    assert np.abs(got[:, 2] - o.z).max() <= 1e-9 * 50.0 / 20


def test_fortran_driver_is_built():
    assert DRIVER.exists()
