"""Graph replay of steady-state steps must give exactly the eagerly launched
result, on every replay, with eager work (gathers, reductions) in between."""
import numpy as np
import pytest

from ch4g import abi, Lattice
import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_DEV])
def test_graph_replay_matches_eager(precision):
    cfg = configs.rb2d(NX=64, NY=32, Ra=1e6, precision=precision)
    eager = Lattice(cfg)
    eager.set_graphs(False)
    eager.initial_conditions()
    graph = Lattice(cfg)
    graph.initial_conditions()
    for rep in range(4):
        n = 70 if rep % 2 == 0 else 64  # 2 graphs + 6 eager steps / exactly 2 graphs
        eager.step(n)
        graph.step(n)
        assert graph.itime == eager.itime
        pe, ge = eager.get_populations()
        pg, gg = graph.get_populations()
        assert np.array_equal(pe, pg) and np.array_equal(ge, gg), rep
        graph.dump_averages()  # eager work between replays


def test_graph_replay_les_matches_eager():
    cfg = configs.poiseuille(8, 16, 8, tau=0.52, g=1e-5)
    cfg.flags |= abi.FLUID_LES | abi.FLUID_LES_LOCAL_STRAIN
    eager = Lattice(cfg)
    eager.set_graphs(False)
    eager.initial_conditions()
    graph = Lattice(cfg)
    graph.initial_conditions()
    for _ in range(3):
        eager.step(65)
        graph.step(65)
        pe, _ = eager.get_populations()
        pg, _ = graph.get_populations()
        assert np.array_equal(pe, pg)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_DEV])
def test_graph_replay_odd_counts_fd_les(precision):
    """Graph replay of a fluid-only channel with the finite-difference LES
    closure (two kernels per step) over step counts that leave the state in
    either buffer."""
    cfg = configs.poiseuille(12, 20, 6, tau=0.52, g=1e-5, precision=precision)
    cfg.flags |= abi.FLUID_LES
    eager = Lattice(cfg)
    eager.set_graphs(False)
    eager.initial_conditions()
    graph = Lattice(cfg)
    graph.initial_conditions()
    for n in (1, 33, 67, 1, 64):
        eager.step(n)
        graph.step(n)
        pe, _ = eager.get_populations()
        pg, _ = graph.get_populations()
        assert np.array_equal(pe, pg), n


@pytest.mark.parametrize("NX,NY,precision,fluid_only", [
    (1024, 704, abi.FP32_DEV, False), (256, 102, abi.FP32_DEV, False), (128, 62, abi.FP32_STORE, False),
    (300, 97, abi.FP32_DEV, False), (40, 30, abi.FP32_DEV, False), (96, 40, abi.FP32_DEV, True)])
def test_two_step_launches_match_single_steps(NX, NY, precision, fluid_only):
    """2-D lattices advance two steps per launch (LDS tile with a one-node
    halo): 1024x704 fluid+T (219 MB, streamed, 64x5 tiles), 256x102
    (cache-resident, non-temporal loads, 64x4 tiles, a partial last tile
    row), 128x62 with double arithmetic, 300x97 (a narrower last tile in x
    and in y), 40x30 (one tile narrower than 64) and a 96x40 fluid-only
    channel.  The same lattice stepped one step per ch4g_step_batch call
    runs single steps only (the batch's single-step slots).  Populations
    must agree bit for bit after odd and even counts and after graph
    replays."""
    from ch4g import step_batch, workloads
    if fluid_only:
        cfg = workloads.channel3d(NX, NY, 1, precision=precision)
    else:
        cfg = configs.rb2d(NX=NX, NY=NY, Ra=1e8 if NY > 500 else 1e6, precision=precision)
    two, one = Lattice(cfg), Lattice(cfg)
    assert two.steps_per_launch() == 2
    for lat in (two, one):
        lat.initial_conditions()
        lat.step(1)
    for n in (7, 10, 70):
        two.step(n)
        for _ in range(n):
            step_batch([one], 1)
        pa, ga = two.get_populations()
        pb, gb = one.get_populations()
        assert np.array_equal(pa, pb) and np.array_equal(ga, gb), n


@pytest.mark.parametrize("shape,precision,fluid_only", [
    ((120, 30, 602), abi.FP32_DEV, False), ((128, 64, 256), abi.FP32_STORE, False),
    ((1024, 1024, 2), abi.FP32_DEV, False), ((192, 96, 120), abi.FP32_DEV, True),
    ((256, 128, 130), abi.FP32_DEV, False)],
    ids=["120x30x602", "128x64x256_arith64", "1024x1024x2", "channel192x96x120", "256x128x130"])
def test_two_step_3d_launches_match_single_steps(shape, precision, fluid_only):
    """3-D lattices of at least 2^21 nodes advance two steps per launch
    (z-marching x-y tiles with an LDS ring of three planes): fluid + T with a
    narrower last tile in x (120 = 64 + 56), a partial last tile row in y (30)
    and a last z range of two planes (602), double arithmetic (64 x 4 tiles),
    NZ = 2 (the z-halo planes are the same plane), a fluid-only channel and a
    lattice streamed from HBM.  The same lattice advanced one step per
    ch4g_step call runs the single-step kernel.  Populations must agree bit
    for bit after odd and even counts and after graph replays."""
    from ch4g import workloads
    NX, NY, NZ = shape
    if fluid_only:
        cfg = workloads.channel3d(NX, NY, NZ, precision=precision)
    else:
        cfg = configs.rb2d(NX=NX, NY=NY, NZ=NZ, Ra=2e6, precision=precision, seed=31)
    print(f"  3-D two-step vs single steps {shape}", flush=True)
    two, one = Lattice(cfg), Lattice(cfg)
    assert two.steps_per_launch() == 2
    for lat in (two, one):
        lat.initial_conditions()
        lat.step(1)
    for n in (7, 34):
        two.step(n)
        for _ in range(n):
            one.step(1)
        pa, ga = two.get_populations()
        pb, gb = one.get_populations()
        assert np.array_equal(pa, pb) and np.array_equal(ga, gb), n
        del pa, ga, pb, gb


def test_small_3d_lattices_step_one_per_launch():
    """Below 2^21 nodes, or with rows that fill less than 90 % of their
    64-wide tiles, 3-D lattices take one step per launch (faster there)."""
    from ch4g import workloads
    for cfg in (workloads.channel3d(96, 96, 96), workloads.channel3d(160, 160, 160),
                configs.rb2d(NX=45, NY=32, NZ=24, Ra=2e6, precision=abi.FP32_DEV)):
        lat = Lattice(cfg)
        assert lat.steps_per_launch() == 1
        lat.close()
