"""The drop-in layer: reference-named functions over the C-ABI, the
define.h / param.in reader, and the main.c-style host ch4g_run."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from ch4g import abi
from ch4g.lattice import PKG_ROOT

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "poiseuille_4x64x4")
DROPIN = os.path.join(PKG_ROOT, "lib", "libch4g_dropin.so")
RUN = os.path.join(PKG_ROOT, "bin", "ch4g_run")

REFERENCE_NAMES = ["initial_conditions", "initialization_forcing", "compute_advection", "add_collision",
                   "build_forcing", "add_forcing", "boundary_and_pbc_conditions_for_streaming", "time_stepping",
                   "hydro_fields", "dump_averages", "measure_time"]


class RunParams(C.Structure):
    _fields_ = [("cfg", abi.Config), ("time_dt", C.c_double), ("time_start", C.c_double),
                ("time_end", C.c_double), ("time_dump_field", C.c_double), ("time_dump_diagn", C.c_double),
                ("output_at_start", C.c_int)]


def test_dropin_exports_reference_names():
    lib = C.CDLL(DROPIN)
    for n in REFERENCE_NAMES + ["ch4g_read_run_params", "ch4g_dropin_init", "ch4g_dropin_finalize"]:
        assert hasattr(lib, n), n


def test_read_run_params_poiseuille():
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    rc = lib.ch4g_read_run_params(os.path.join(FIX, "define.h").encode(), os.path.join(FIX, "param.in").encode(),
                                  C.byref(rp))
    assert rc == 0
    c = rp.cfg
    assert (c.NX, c.NY, c.NZ) == (4, 64, 4)
    assert c.flags == abi.LB_FLUID | abi.FLUID_FORCING_POISEUILLE | abi.FLUID_BC_Y
    assert c.tau_u == 1.0 and c.Amp_x == 0.003072 and c.seed == 1
    assert rp.time_end == 1000.0 and rp.time_dump_diagn == 250.0 and rp.output_at_start == 1


def test_read_run_params_rejects_unimplemented_switch(tmp_path):
    d = tmp_path / "define.h"
    d.write_text(open(os.path.join(FIX, "define.h")).read() + "#define LB_FLUID_BC_Y_P_SLIP\n")
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(str(d).encode(), os.path.join(FIX, "param.in").encode(), C.byref(rp)) == -1


def test_read_run_params_missing_key(tmp_path):
    p = tmp_path / "param.in"
    p.write_text("\n".join(l for l in open(os.path.join(FIX, "param.in")) if not l.startswith("tau_u")))
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(os.path.join(FIX, "define.h").encode(), str(p).encode(), C.byref(rp)) == -1


@pytest.mark.gpu
def test_ch4g_run_matches_oracle(tmp_path):
    """ch4g_run on the Poiseuille fixture writes velocity_averages.dat at
    t = 0, 250, ..., 1000 equal to the oracle's averages at those steps."""
    from oracle import Oracle
    import configs
    for f in ("define.h", "param.in"):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    r = subprocess.run([RUN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "The End" in r.stdout
    rows = np.loadtxt(tmp_path / "velocity_averages.dat")
    assert rows.shape == (5, 11)
    assert (tmp_path / "run_time.dat").exists()
    cfg = configs.poiseuille(4, 64, 4)
    cfg.Amp_x = 0.003072
    o = Oracle(cfg)
    o.initial_conditions()
    for k, t in enumerate(rows[:, 0]):
        if k:
            o.step(250)
        a = o.dump_averages()
        assert t == 250 * k
        np.testing.assert_allclose(rows[k, 3], a["ux"], rtol=1e-6)  # %e output: 7 significant digits
        np.testing.assert_allclose(rows[k, 1], a["ene"], rtol=1e-6, atol=1e-30)
