"""ch4g_dump_averages_noalloc against ch4g_dump_averages, through the C-ABI.

The allocation-free entry point computes the averages from caller-owned
device field buffers with the block partials in the library's static device
array.  It must give the same averages as ch4g_dump_averages (bit-identical:
same gather, same reduction grid and order), leave the caller's buffers
holding the hydro fields, and allocate nothing (the lattice's device byte
count is unchanged by the call).  Two calls in a row, the second on a smaller
lattice, check that the shared array carries nothing over between calls.
"""
import numpy as np
import pytest
import torch

from ch4g import abi, Lattice
import configs

pytestmark = pytest.mark.gpu


def _buffers(lat, fluid, thermal):
    n = lat.cfg.nodes()
    dev = torch.device("cuda:0")
    d = torch.full((n,), np.nan, dtype=torch.float64, device=dev) if fluid else None
    u = torch.full((3 * n,), np.nan, dtype=torch.float64, device=dev) if fluid else None
    t = torch.full((n,), np.nan, dtype=torch.float64, device=dev) if thermal else None
    torch.cuda.synchronize()
    return d, u, t


def _ptr(x):
    return None if x is None else x.data_ptr()


def _check(lat, fluid, thermal, label):
    d, u, t = _buffers(lat, fluid, thermal)
    before = lat.device_bytes
    got = lat.dump_averages_noalloc(_ptr(d), _ptr(u), _ptr(t))
    assert lat.device_bytes == before, "the allocation-free entry point allocated device memory"
    ref = lat.dump_averages()
    for k, v in ref.items():
        assert got[k] == v or (np.isnan(got[k]) and np.isnan(v)), (label, k, got[k], v)
    h = lat.hydro_fields()
    if fluid:
        assert np.array_equal(d.cpu().numpy(), h["dens"].ravel()), label
        assert np.array_equal(u.cpu().numpy(), h["u"].reshape(-1)), label
    if thermal:
        assert np.array_equal(t.cpu().numpy(), h["t"].ravel()), label
    print(f"  {label}: nuy {got['nuy']:.12e} ene {got['ene']:.12e} bit-identical", flush=True)


def test_noalloc_averages_match_two_calls_second_smaller():
    # first call: a 3-D Rayleigh-Benard lattice in FP32-deviation storage
    cfg = configs.rb2d(NX=96, NY=64, NZ=48, Ra=2e6, precision=abi.FP32_DEV, seed=99)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(40)
    _check(lat, True, True, "rb3d 96x64x48 fp32dev")
    lat.close()
    # second call: a smaller 2-D lattice in FP64, then a fluid-only channel
    cfg = configs.rb2d(NX=64, NY=32, Ra=1e6, seed=5)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(25)
    _check(lat, True, True, "rb2d 64x32 fp64")
    lat.close()
    cfg = configs.poiseuille(4, 64, 4)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(30)
    _check(lat, True, False, "poiseuille 4x64x4 fp64 (no temperature buffer)")
    lat.close()


def test_noalloc_rejects_missing_buffers():
    cfg = configs.rb2d(NX=32, NY=16, Ra=1e5)
    lat = Lattice(cfg)
    lat.initial_conditions()
    d, u, _ = _buffers(lat, True, False)
    with pytest.raises(Exception, match="thermal lattice needs"):
        lat.dump_averages_noalloc(_ptr(d), _ptr(u), None)
    with pytest.raises(Exception, match="fluid lattice needs"):
        lat.dump_averages_noalloc(None, None, None)
    lat.close()
