"""Pins of the CPU oracle (the checker of every GPU parity test).

Each check is chosen so that a plausible mistake fails it: a wrong sign in the DFT, a dropped
1/n^2, R^2 replaced by n^2 in the Gaussian exponent (box.cpp:636-637), a transposed spectrum
index, a broken comparison / redistancing / contour would each move one of these numbers.
"""
import math
import os

import numpy as np
import pytest

from gragles2d_amd import grainhdl as G

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 30, 37, 64, 90, 127, 128, 256])
@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_matches_numpy(oracle, n, sign):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    got = oracle.dft(x, sign)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1, n)


def test_convolution_matches_numpy_golden(oracle):
    g = np.load(os.path.join(HERE, "golden", "conv_golden.npz"))
    dt, R = float(g["dt"]), int(g["R"])
    for key in g.files:
        if not key.startswith("phi_"):
            continue
        n = int(key[4:])
        got = oracle.convolve(g[key], 2, dt, R)
        assert np.max(np.abs(got - g[f"conv_{n}"])) < 1e-15, n


@pytest.mark.parametrize("n,kx,ky", [(40, 3, 0), (40, 0, 5), (37, 2, 7), (64, 9, 4)])
def test_single_cosine_mode_is_scaled_by_its_symbol(oracle, n, kx, ky):
    # phi = cos(2 pi (kx x + ky y)/n) is an eigenfunction: out = exp(-c (kx^2+ky^2)) phi with
    # c = sqrt(2/pi) 4 dt R^2 pi^2 / n^2 -- R, not n, in the numerator (box.cpp:636-637,682)
    R, dt = 894, 0.8 / 894**2 * 20 / 10
    y, x = np.mgrid[0:n, 0:n]
    phi = np.cos(2 * np.pi * (kx * x + ky * y) / n)
    c = math.sqrt(2 / math.pi) * 4 * dt * R * R * math.pi**2 / (n * n)
    got = oracle.convolve(phi, 2, dt, R)
    assert np.max(np.abs(got - math.exp(-c * (kx * kx + ky * ky)) * phi)) < 1e-13


def test_convolution_dc_gain_and_self_adjointness(oracle):
    rng = np.random.default_rng(5)
    R, dt = 447, 0.8 / 447**2 * 20 / 10
    n = 45
    assert np.allclose(oracle.convolve(np.full((n, n), 0.25), 2, dt, R), 0.25, atol=1e-15)
    a, b = rng.normal(size=(n, n)), rng.normal(size=(n, n))
    lhs = np.sum(oracle.convolve(a, 2, dt, R) * b)
    rhs = np.sum(a * oracle.convolve(b, 2, dt, R))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_laplace_mode_symbol(oracle):
    # mode 0: G = 1/(1 + dt 2(2 - cos kx - cos ky) R^2)/n^2 (box.cpp:648-659)
    n, kx, ky, R, dt = 32, 3, 1, 200, 1e-5
    y, x = np.mgrid[0:n, 0:n]
    phi = np.cos(2 * np.pi * (kx * x + ky * y) / n)
    k = 2 * np.pi / n
    G = 1.0 / (1.0 + dt * 2 * (2 - math.cos(k * kx) - math.cos(k * ky)) * R * R)
    assert np.max(np.abs(oracle.convolve(phi, 0, dt, R) - G * phi)) < 1e-13


def _run_oracle(net, R, gb, steps, ppg=20):
    s = G.default_settings(number_of_particles=len(net), number_of_points_per_grain=ppg,
                           domain_border_size=gb, convolution_mode=2, is_isotropic_network=1,
                           analysis_timestep=1, microstructure_gen_mode=0)
    d = G.derive(s)
    from oracle.pyoracle import Oracle
    o = Oracle(R if R else d.real_domain_size, gb, d.dt, len(net), analysis_timestep=1)
    for i in range(len(net)):
        gid, pts, q = net.grain(i)
        o.add_grain(gid, pts, q)
    o.init()
    hist = []
    for _ in range(steps):
        assert o.step() == 0
        hist.append([o.info(g).volume if o.info(g).exists else 0.0 for g in range(1, len(net) + 1)])
    return o, np.array(hist)


def test_straight_boundary_is_stationary():
    # two grains split by x = 0.5: the interface has zero curvature and stays put; both grains
    # lose the same area where they meet the domain wall (grain 0, boundaryCondition)
    left = np.array([[0, 0], [0.5, 0], [0.5, 1], [0, 1], [0, 0]], float)
    right = np.array([[0.5, 0], [1, 0], [1, 1], [0.5, 1], [0.5, 0]], float)
    net = G.Network.from_polygons([left, right])
    gb = 7
    R = int(math.sqrt(2) * 40)
    o, hist = _run_oracle(net, 0, gb, 6, ppg=40)
    assert np.all(np.abs(hist[:, 0] - hist[:, 1]) < 1e-3)
    assert np.all(np.diff(hist[:, 0]) <= 0) and hist[-1, 0] > 0.48
    c = o.contour(1)
    mid = c[np.abs(c[:, 1] - (0.5 * R + gb)) < 10]
    xs = mid[mid[:, 0] > R / 4 + gb, 0]
    assert len(xs) > 5 and np.all(np.abs(xs - (0.5 * R + gb)) < 0.05), xs


def _disk_in_quadrants(r=0.22, npts=64):
    t = np.linspace(0, 2 * np.pi, npts + 1)
    disk = np.stack([0.5 + r * np.cos(t), 0.5 + r * np.sin(t)], 1)
    disk[-1] = disk[0]
    quads = []
    for qx, qy, a0 in ((1, 1, 0.0), (0, 1, 0.5 * np.pi), (0, 0, np.pi), (1, 0, 1.5 * np.pi)):
        arc = np.linspace(a0, a0 + 0.5 * np.pi, npts // 4 + 1)
        pts = [[0.5 + r * np.cos(a), 0.5 + r * np.sin(a)] for a in arc]
        corner_far = [qx * 1.0, qy * 1.0]
        # from the arc end, go out along the axis, around the far corner, back along the other axis
        end, start = pts[-1], pts[0]
        ax_end = [0.5 + (end[0] - 0.5) / r * 0.5 if abs(end[0] - 0.5) > 1e-12 else 0.5,
                  0.5 + (end[1] - 0.5) / r * 0.5 if abs(end[1] - 0.5) > 1e-12 else 0.5]
        ax_start = [0.5 + (start[0] - 0.5) / r * 0.5 if abs(start[0] - 0.5) > 1e-12 else 0.5,
                    0.5 + (start[1] - 0.5) / r * 0.5 if abs(start[1] - 0.5) > 1e-12 else 0.5]
        poly = pts + [ax_end, corner_far, ax_start, pts[0]]
        quads.append(np.array(poly))
    return disk, quads


def test_four_sided_grain_shrinks_at_constant_rate():
    # von Neumann-Mullins: a grain with four neighbours shrinks at a constant dA/dt
    disk, quads = _disk_in_quadrants()
    net = G.Network.from_polygons([disk] + quads)
    o, hist = _run_oracle(net, 0, 7, 14, ppg=40)
    total = hist.sum(1)
    assert np.all(np.abs(total - total[0]) < 5e-3)            # the tiling stays a tiling
    a = hist[4:, 0]
    rate = np.diff(a)
    assert np.all(rate < 0)
    assert np.std(rate) < 0.2 * abs(np.mean(rate)), rate
    assert len(o.neighbours(1)[0]) == 4


def test_voronoi_network_invariants():
    net = G.Network.generate(60, 1000)
    o, hist = _run_oracle(net, 0, 7, 8)
    # area is conserved up to the domain wall region
    assert abs(hist[-1].sum() - 1.0) < 0.05, hist[-1].sum()
    # neighbour relation is (almost) symmetric and lengths agree to O(h)
    asym = 0
    pairs = 0
    for g in range(1, 61):
        if not o.info(g).exists:
            continue
        ids, lens = o.neighbours(g)
        for nb, ln in zip(ids, lens):
            if nb == 0:
                continue
            pairs += 1
            ids2, lens2 = o.neighbours(int(nb))
            if g not in ids2:
                asym += 1
    assert pairs > 100 and asym <= 0.05 * pairs
    # mean number of grain neighbours of a 2-D network is close to 6
    counts = [len([x for x in o.neighbours(g)[0] if x != 0]) for g in range(1, 61) if o.info(g).exists and not o.info(g).intersects_boundary]
    assert 4.5 < np.mean(counts) < 7.5
