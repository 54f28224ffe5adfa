"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, and the ctypes mirror matches the C struct layout."""
import ctypes as C
import os
import re
import subprocess

import pytest

from ch4g import abi
from ch4g.lattice import EXPORTS, INCLUDE_DIR, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ch4g_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "run __graft_entry__.build() first"
    lib = C.CDLL(LIB_PATH)
    decl = declared_functions(os.path.join(INCLUDE_DIR, "ch4g.h"))
    assert decl, "no declarations parsed"
    missing = [f for f in decl if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(EXPORTS) == decl


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "ch4g.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(ch4g_config), offsetof(ch4g_config, tau_u),"
        " offsetof(ch4g_config, T0), sizeof(ch4g_output), offsetof(ch4g_output, nuy));"
        " printf(\"%zu %zu %zu %zu\\n\", sizeof(ch4g_particle_config), offsetof(ch4g_particle_config, tau_drag_min),"
        " sizeof(ch4g_particle), offsetof(ch4g_particle, Dt_uy));"
        " printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(ch4g_job), offsetof(ch4g_job, out), offsetof(ch4g_job, status),"
        " sizeof(ch4g_serve_params), offsetof(ch4g_serve_params, use_graphs)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", INCLUDE_DIR, str(src), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    want = [C.sizeof(abi.Config), abi.Config.tau_u.offset, abi.Config.T0.offset, C.sizeof(abi.Output),
            abi.Output.nuy.offset, C.sizeof(abi.ParticleConfig), abi.ParticleConfig.tau_drag_min.offset,
            C.sizeof(abi.Particle), abi.Particle.Dt_uy.offset, C.sizeof(abi.Job), abi.Job.out.offset,
            abi.Job.status.offset, C.sizeof(abi.ServeParams), abi.ServeParams.use_graphs.offset]
    assert abi.particle_dtype().itemsize == C.sizeof(abi.Particle)
    assert got == want


def test_footprint_and_errors_without_gpu():
    lib = C.CDLL(LIB_PATH)
    lib.ch4g_footprint_bytes.restype = C.c_int64
    # two buffers of 19 fluid and 19 temperature planes plus the lag scalar;
    # planes whose byte size is a multiple of 64 KiB are padded by 4160
    # elements (512x256 floats: 512 KiB planes), others are not (500x256)
    cfg = abi.make_config(512, 256, 1, abi.LB_FLUID | abi.LB_TEMPERATURE, precision=abi.FP32_DEV)
    n = 512 * 256
    assert lib.ch4g_footprint_bytes(C.byref(cfg)) == 4 * 19 * (n + 4160) * 4 + n * 4
    cfg = abi.make_config(500, 256, 1, abi.LB_FLUID | abi.LB_TEMPERATURE, precision=abi.FP32_DEV)
    n = 500 * 256
    assert lib.ch4g_footprint_bytes(C.byref(cfg)) == 4 * 19 * n * 4 + n * 4
    cfg = abi.make_config(512, 256, 1, abi.LB_FLUID, precision=abi.FP64)
    n = 512 * 256
    assert lib.ch4g_footprint_bytes(C.byref(cfg)) == 2 * 19 * (n + 4160) * 8
    if lib.ch4g_device_count() == 0:
        h = C.c_void_p()
        rc = lib.ch4g_create(C.byref(cfg), C.byref(h))
        assert rc < 0
        lib.ch4g_last_error.restype = C.c_char_p
        assert b"device" in lib.ch4g_last_error()


def test_invalid_configs_rejected_before_device_use():
    lib = C.CDLL(LIB_PATH)
    lib.ch4g_last_error.restype = C.c_char_p
    h = C.c_void_p()
    # buoyancy without the temperature populations
    cfg = abi.make_config(8, 8, 8, abi.LB_FLUID | abi.TEMPERATURE_BUOYANCY, precision=abi.FP32_DEV)
    assert lib.ch4g_create(C.byref(cfg), C.byref(h)) == -1
    assert b"LB_TEMPERATURE" in lib.ch4g_last_error()
    # y walls on a lattice of fewer than three rows
    cfg = abi.make_config(8, 2, 8, abi.LB_FLUID | abi.FLUID_BC_Y, precision=abi.FP32_DEV)
    assert lib.ch4g_create(C.byref(cfg), C.byref(h)) == -1
    assert b"NY >= 3" in lib.ch4g_last_error()


def test_particle_record_size_matches_the_library():
    """ch4g_particle_size reports the record stride the library was built
    with; the Python binding refuses a library whose stride differs."""
    lib = C.CDLL(LIB_PATH)
    lib.ch4g_particle_size.restype = C.c_int64
    assert lib.ch4g_particle_size() == C.sizeof(abi.Particle) == 8 * len(abi.PARTICLE_FIELDS)
