"""Benchmark of the batched per-grain level-set timestep on one MI355X.

One step = one pass of the run_sim loop body over every live grain of the workload (default
batch-100k: ~100k grains of a lognormal power diagram, R = 6324, box sides ~16-110, FP64).
Inputs are resident in HBM before the timed region; the timed region is exactly --steps steps
between two device synchronisations.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batch-100k")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--pergrain-steps", type=int, default=1, help="steps timed with one launch per grain (0 = skip)")
    ap.add_argument("--cpu-grains", type=int, default=3000, help="grains of the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--persistent", type=int, default=0)
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)
    from gragles2d_amd import _native as N
    from gragles2d_amd import grainhdl as G
    from gragles2d_amd.configs import WORKLOADS

    w = WORKLOADS[a.workload]
    prec = N.PREC_FP64 if a.precision == "fp64" else N.PREC_FP32
    S = 8 if prec == N.PREC_FP64 else 4
    net = G.Network.generate(w.grains, w.seed, w.lognormal_sigma)
    kw = w.settings_kw()
    kw["number_of_particles"] = len(net)
    s = G.default_settings(**kw)
    hdl = G.GrainHandler(s, precision=prec, batching=N.BATCH_FUSED, persistent=a.persistent)
    hdl.set_simulation_parameter(net)
    hdl.step(a.warmup)
    hdl.sync()
    st0 = hdl.stats()
    t0 = time.perf_counter()
    hdl.step(a.steps)
    hdl.sync()
    t1 = time.perf_counter()
    st1 = hdl.stats()
    upd = st1.grid_point_updates - st0.grid_point_updates
    sec = t1 - t0
    value = upd / sec
    phases = {k: (getattr(st1, k) - getattr(st0, k)) / a.steps for k in
              ("convolution_ms", "comparison_ms", "levelset_ms")}
    # algorithmic bytes per step of the two big phases (the library's device counters)
    conv_b = (st1.conv_bytes - st0.conv_bytes) / a.steps
    level_b = (st1.algorithmic_bytes - st0.algorithmic_bytes) / a.steps - conv_b
    if phases["levelset_ms"] >= phases["convolution_ms"]:
        kern, byts, ms = "k_level (fused comparison / contour / neighbours / redistancing)", level_b, phases["levelset_ms"]
    else:
        kern, byts, ms = "k_conv_dense (convolution)", conv_b, phases["convolution_ms"]
    achieved = byts / (ms * 1e-3)
    traffic = measured_traffic(w.name, a.precision, kern)
    out = {
        "metric": "grid-point updates/s (batched timestep)",
        "value": value,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == N.PREC_FP64 else "f32",
        "data": "synthetic (seeded lognormal power diagram)",
        "config": {"workload": w.name, "grains": len(net), "grid": G.derive(s).real_domain_size,
                   "live_grains_last_step": st1.live_grains, "max_box": st1.max_box,
                   "grid_points_per_step": upd / a.steps},
        "phases_ms_per_step": phases,
        "algorithmic_bytes_per_step": {"convolution": conv_b, "level": level_b},
        "roofline": {"kernel": kern, "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic},
        "step_roofline_frac": (conv_b + level_b) / (sec / a.steps) / HBM_PEAK,
    }
    hdl.close()
    if a.pergrain_steps > 0:
        pg = G.GrainHandler(s, precision=prec, batching=N.BATCH_PERGRAIN)
        pg.set_simulation_parameter(net)
        pg.step(1)
        pg.sync()
        u0 = pg.stats().grid_point_updates
        t0 = time.perf_counter()
        pg.step(a.pergrain_steps)
        pg.sync()
        pgs = time.perf_counter() - t0
        u2 = pg.stats().grid_point_updates - u0
        out["pergrain_value"] = u2 / pgs
        out["batched_over_pergrain"] = value / (u2 / pgs)
        pg.close()
    if a.cpu_grains > 0:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_grains, a.cpu_steps)
    print(json.dumps(out), flush=True)


def measured_traffic(workload, precision, kernel):
    """HBM bytes per step of the dominant phase from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, scaled by what the counters read for a copy of known size);
    None when no summary exists for this workload and precision."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload.replace('-', '')}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    t_name = "double" if precision == "fp64" else "float"
    name = f"ggx::k_level<{t_name}, true>" if kernel.startswith("k_level") else f"ggx::k_conv_dense<{t_name}>"
    e = t.get("per_step", {}).get(name)
    return None if e is None else e["hbm_bytes"]


def cpu_baseline(w, ngrains, steps):
    """The oracle (C++/OpenMP FP64 restatement) on a bounded sample of the same workload."""
    from gragles2d_amd import grainhdl as G
    from oracle.pyoracle import Oracle

    threads = min(16, os.cpu_count() or 1)
    net = G.Network.generate(ngrains, w.seed, w.lognormal_sigma)
    kw = w.settings_kw()
    kw["number_of_particles"] = len(net)
    d = G.derive(G.default_settings(**kw))
    o = Oracle(d.real_domain_size, w.border, d.dt, len(net), analysis_timestep=kw["analysis_timestep"], nthreads=threads)
    for i in range(len(net)):
        gid, pts, q = net.grain(i)
        o.add_grain(gid, pts, q)
    o.init()
    o.step()
    upd, sec = 0, 0.0
    for _ in range(steps):
        for g in range(1, len(net) + 1):
            inf = o.info(g)
            if inf.present and inf.exists:
                upd += (inf.xmax - inf.xmin) ** 2
        t0 = time.perf_counter()
        o.step()
        sec += time.perf_counter() - t0
    return {"value": upd / sec, "unit": "grid-point updates/s", "cores": threads, "kind": "port",
            "sample": f"{len(net)}-grain network of the same workload (ppg {w.points_per_grain}, border {w.border}, "
                      f"sigma {w.lognormal_sigma}), {steps} steps after one warm-up step"}


if __name__ == "__main__":
    main()
