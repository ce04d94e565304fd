"""Shared helpers of the GPU parity tests: run libggx and the oracle on the same network."""
import numpy as np

from gragles2d_amd import _native as N
from gragles2d_amd import grainhdl as G


def settings_for(ngrains, ppg=20, gb=7, **kw):
    return G.default_settings(number_of_particles=ngrains, number_of_points_per_grain=ppg,
                              domain_border_size=gb, convolution_mode=2, is_isotropic_network=1,
                              analysis_timestep=1, microstructure_gen_mode=0, **kw)


def make_pair(net, ppg=20, gb=7, precision=N.PREC_FP64, batching=N.BATCH_FUSED, oracle_threads=0, **opts):
    from oracle.pyoracle import Oracle
    s = settings_for(len(net), ppg, gb)
    d = G.derive(s)
    gpu = G.GrainHandler(s, precision=precision, batching=batching, **opts)
    gpu.set_simulation_parameter(net)
    o = Oracle(d.real_domain_size, gb, d.dt, len(net), analysis_timestep=1, nthreads=oracle_threads)
    for i in range(len(net)):
        gid, pts, q = net.grain(i)
        o.add_grain(gid, pts, q)
    o.init()
    return gpu, o


def tie_mask(gpu_or_oracle_fields):
    return None


def compare_step(gpu, o, ids, phi_tol=1e-10, len_tol=1e-9, check_idlocal=True):
    """Returns a dict of discrepancies after a step (empty dict = parity)."""
    bad = {}
    ties = {}
    worst = 0.0
    idl_mis = 0
    gi_all = gpu.info_all()
    for g in ids:
        gi, oi = gi_all[g], o.info(g)
        if (gi.present, gi.present and gi.exists) != (oi.present, oi.present and oi.exists):
            bad.setdefault("existence", []).append(g)
            continue
        if not (oi.present and oi.exists):
            continue
        if (gi.xmin, gi.xmax, gi.ymin, gi.ymax) != (oi.xmin, oi.xmax, oi.ymin, oi.ymax):
            bad.setdefault("box", []).append(g)
            continue
        if (gi.motion_regular, gi.intersects_boundary) != (oi.motion_regular, oi.intersects_boundary):
            bad.setdefault("flags", []).append(g)
        if gi.contour_len != oi.contour_len:
            ties["contour_points"] = ties.get("contour_points", 0) + 1
        for key in ("volume", "perimeter", "energy"):
            a, b = getattr(gi, key), getattr(oi, key)
            if abs(a - b) > len_tol * max(1.0, abs(b)):
                bad.setdefault(key, []).append(g)
        d = np.max(np.abs(gpu.phi(g) - o.phi(g)))
        worst = max(worst, d)
        if d > phi_tol:
            bad.setdefault("phi", []).append((g, d))
        ni, li = gpu.neighbours(g)
        no, lo = o.neighbours(g)
        if not np.array_equal(ni, no) or (len(li) and np.max(np.abs(li - lo)) > len_tol):
            bad.setdefault("neighbours", []).append(g)
        if not np.array_equal(gpu.compare_list(g), o.compare_list(g)):
            bad.setdefault("compare_list", []).append(g)
        if check_idlocal:
            a, b = gpu.idlocal(g), o.idlocal(g)
            if a.shape != b.shape:
                bad.setdefault("idlocal_shape", []).append(g)
            else:
                idl_mis += int(np.sum(a != b))
    # IDLocal pixels and contour point counts can differ only where the oracle's FFT and the
    # GPU's circulant round an exact tie differently (e.g. a grain edge lying on the domain wall
    # makes the convolved field exactly antisymmetric there); they are reported, the parity bar
    # is the fields, boxes, existence and neighbour lists
    if idl_mis:
        ties["idlocal_pixels"] = idl_mis
    bad["_max_phi_err"] = worst
    bad["_ties"] = ties
    return bad


def failures(report):
    return {k: v for k, v in report.items() if not k.startswith("_")}
