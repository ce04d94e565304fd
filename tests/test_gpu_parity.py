"""GPU parity: libggx (HIP, through the C-ABI) against the CPU oracle on the same seeded networks.

Bar: FP64 signed-distance fields within 1e-10 max-abs per grain after every
step, identical surviving-grain sets, identical box bounds, identical neighbour lists (IDs in
first-encounter order; boundary lengths within 1e-9), identical comparison lists; FP32 grain areas
within 1e-5 relative.
"""
import sys

import numpy as np
import pytest

from gragles2d_amd import _native as N
from gragles2d_amd import grainhdl as G
from gragles2d_amd.configs import WORKLOADS

from .gpu_helpers import compare_step, failures, make_pair

pytestmark = pytest.mark.gpu


def _run(net, steps, ppg=20, gb=7, **kw):
    gpu, o = make_pair(net, ppg=ppg, gb=gb, **kw)
    ids = range(1, len(net) + 1)
    worst = 0.0
    for t in range(steps):
        gpu.step(1)
        assert o.step() == 0
        rep = compare_step(gpu, o, ids)
        worst = max(worst, rep["_max_phi_err"])
        print(f"step {t}: grains {gpu.current_grains()}/{o.current_grains()} max|dphi| {rep['_max_phi_err']:.2e}", flush=True)
        assert not failures(rep), (t, failures(rep))
        assert gpu.current_grains() == o.current_grains()
    return gpu, o, worst


def test_plain100_fp64_every_step():
    """Workload plain-100 (100 grains, ~200x200, 20 timesteps -> 21 passes of the loop)."""
    w = WORKLOADS["plain-100"]
    net = G.Network.generate(w.grains, w.seed, w.lognormal_sigma)
    _, _, worst = _run(net, w.steps + 1, w.points_per_grain, w.border)
    assert worst <= 1e-10


@pytest.mark.parametrize("seed,gb,sigma", [(7, 10, 0.0), (99, 7, 0.35), (5, 20, 0.0)])
def test_seeds_margins_and_lognormal(seed, gb, sigma):
    net = G.Network.generate(150, seed, sigma)
    _run(net, 8, 20, gb)


def test_iso2k_fp64_first_steps():
    """Workload iso-2k (2000 grains, ~900x900, box sides 40-90) for its first 12 passes."""
    w = WORKLOADS["iso-2k"]
    net = G.Network.generate(w.grains, w.seed, w.lognormal_sigma)
    _run(net, 12, w.points_per_grain, w.border, oracle_threads=16)


def test_tiny_and_edge_networks():
    # one grain filling the domain; two grains; a grain smaller than the margin
    one = [np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0, 0]], float)]
    _run(G.Network.from_polygons(one), 3)
    two = [np.array([[0, 0], [0.5, 0], [0.5, 1], [0, 1], [0, 0]], float),
           np.array([[0.5, 0], [1, 0], [1, 1], [0.5, 1], [0.5, 0]], float)]
    _run(G.Network.from_polygons(two), 4)
    tiny = [np.array([[0.45, 0.45], [0.47, 0.45], [0.47, 0.47], [0.45, 0.47], [0.45, 0.45]]),
            np.array([[0, 0], [1, 0], [1, 0.45], [0.45, 0.45], [0.45, 0.47], [0.47, 0.47], [0.47, 0.45],
                      [1, 0.45], [1, 1], [0, 1], [0, 0]])]
    _run(G.Network.from_polygons(tiny), 5, ppg=40)


def test_large_boxes_strip_convolution_and_global_tiles():
    """Boxes above 128 points: the strip convolution, the level kernel with C and the output in
    global memory, and the lane-0 walk over C (6 grains, 60 points per grain, margin 40)."""
    net = G.Network.generate(6, 3)
    _run(net, 3, ppg=60, gb=40)


def test_pergrain_launches_match_batched_bitwise():
    net = G.Network.generate(60, 11)
    s = G.default_settings(number_of_particles=len(net), number_of_points_per_grain=20, domain_border_size=7,
                           convolution_mode=2, is_isotropic_network=1, analysis_timestep=1)
    a = G.GrainHandler(s, batching=N.BATCH_FUSED)
    b = G.GrainHandler(s, batching=N.BATCH_PERGRAIN)
    a.set_simulation_parameter(net)
    b.set_simulation_parameter(net)
    for _ in range(4):
        a.step(1)
        b.step(1)
    for g in range(1, len(net) + 1):
        ia, ib = a.info(g), b.info(g)
        assert (ia.present, ia.exists, ia.xmin, ia.ymin, ia.xmax) == (ib.present, ib.exists, ib.xmin, ib.ymin, ib.xmax)
        if ia.present and ia.exists:
            assert np.array_equal(a.phi(g), b.phi(g))


def test_fp32_areas_within_1e5():
    w = WORKLOADS["plain-100"]
    net = G.Network.generate(w.grains, w.seed)
    gpu, o = make_pair(net, precision=N.PREC_FP32)
    for _ in range(w.steps + 1):
        gpu.step(1)
        o.step()
    rel = []
    for g in range(1, len(net) + 1):
        gi, oi = gpu.info(g), o.info(g)
        assert (gi.present and gi.exists) == (oi.present and oi.exists), g
        if oi.present and oi.exists:
            rel.append(abs(gi.volume - oi.volume) / oi.volume)
    print("fp32 max rel area error", max(rel))
    assert max(rel) <= 1e-5


def test_persistent_launches_match_one_workgroup_per_grain_bitwise():
    """The fixed-grid launch mode (grains pulled from the device work lists) gives the same bits."""
    net = G.Network.generate(120, 21, 0.35)
    s = G.default_settings(number_of_particles=len(net), number_of_points_per_grain=20, domain_border_size=7,
                           convolution_mode=2, is_isotropic_network=1, analysis_timestep=1)
    a = G.GrainHandler(s, batching=N.BATCH_FUSED)
    b = G.GrainHandler(s, batching=N.BATCH_FUSED, persistent=1)
    a.set_simulation_parameter(net)
    b.set_simulation_parameter(net)
    for _ in range(5):
        a.step(1)
        b.step(1)
    for g in range(1, len(net) + 1):
        ia, ib = a.info(g), b.info(g)
        assert (ia.present, ia.exists, ia.xmin, ia.ymin, ia.xmax) == (ib.present, ib.exists, ib.xmin, ib.ymin, ib.xmax)
        if ia.present and ia.exists:
            assert np.array_equal(a.phi(g), b.phi(g))
            assert list(a.neighbours(g)[0]) == list(b.neighbours(g)[0])
