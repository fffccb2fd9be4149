#!/usr/bin/env python
"""Energy-exchange throughput on one MI355X.

Workload: the street-canyon scene (100x20 m ground, two 100x15 m facades
with 10 recesses, 1 m patches: 5,122 patches, 32 BRDF directions, 8 bands,
4 kHz bins over 1.5 s = 6,000 bins), one source.  Setup (untimed, reported):
bake (visibility, form factors, direction map), source injection, exchange
plan.  A step is one reflection order of the exchange over the whole scene;
``--warmup`` orders run first, then ``--steps`` orders are timed.

Metric: patch-pair.bin updates per second, U = sum over directed visible
pairs (i->j) of B * max(0, T - n_ij) per order (the reference's update
count, RadiosityFast.py:1137-1143).

Prints one JSON line (see ``--help``).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def _scene(name):
    from sparrowpy_amd import scenes
    if name == "canyon":
        return scenes.street_canyon()
    if name == "shoebox":
        return scenes.directional_shoebox()
    if name == "small":
        return scenes.small_shoebox()
    raise ValueError(name)


def _cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(plan, D, T, seconds, seed=0):
    """Time the FP64 oracle's exchange order on a bounded sample.

    Sample: a seeded subset of destination patches with their full source
    lists, band 0 only, the histogram truncated to ``T_s`` bins, all host
    threads the process may use.  Returns (updates/s, threads, T_s, n_dest).
    """
    import oracle as orc
    import torch
    orc.build()
    lib = orc._load()
    nthreads = int(os.environ.get("OMP_NUM_THREADS",
                                  min(os.cpu_count() or 1, 16)))
    row_ptr = plan.row_ptr.cpu().numpy()
    src = plan.src.cpu().numpy()
    lag = plan.lag.cpu().numpy()
    dirs = plan.dir.cpu().numpy()
    P = plan.P
    T_s = min(T, 2400)
    rng = np.random.default_rng(seed)
    prev = np.full((P, D, 1, T_s), 1e-3)
    cur = np.zeros_like(prev)
    order = rng.permutation(P)

    def run(dests):
        rp = [0]
        s_, l_, d_, c_ = [], [], [], []
        for j in dests:
            a, b = row_ptr[j], row_ptr[j + 1]
            s_.append(src[a:b])
            l_.append(lag[a:b])
            d_.append(dirs[a:b])
            rp.append(rp[-1] + (b - a))
        s_ = np.concatenate(s_).astype(np.int32)
        l_ = np.concatenate(l_).astype(np.int32)
        d_ = np.concatenate(d_).astype(np.int32)
        coef = plan.coef[torch.as_tensor(np.concatenate(
            [np.arange(row_ptr[j], row_ptr[j + 1]) for j in dests]),
            device=plan.coef.device), 0, :D].double().cpu().numpy()
        # destinations re-indexed 0..n-1 in a compact cur buffer
        n = len(dests)
        cur_s = np.zeros((n, D, 1, T_s))
        rp = np.asarray(rp, np.int64)
        upd = int(np.sum(np.maximum(0, T_s - l_.astype(np.int64))))
        t0 = time.perf_counter()
        lib.orc_exchange_order(T_s, P, D, 1, prev, cur_s, rp, s_, l_, d_,
                               np.ascontiguousarray(coef), 0, n, nthreads)
        return upd, time.perf_counter() - t0

    # calibrate on a few destinations, then size the sample to `seconds`
    upd, dt = run(order[:8])
    per_dest = max(dt / 8, 1e-6)
    n_dest = int(min(P, max(8, seconds / per_dest)))
    upd, dt = run(order[:n_dest])
    return upd / dt, nthreads, T_s, n_dest


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="canyon",
                    choices=["canyon", "shoebox", "small"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not "
              "supported (the exchange shards by band with no collectives; "
              "multi-GPU runs are not part of this build)", file=sys.stderr)
        return 2

    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, ops
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData

    sc = _scene(args.config)
    torch.cuda.init()
    dev = torch.device("cuda", 0)
    t = {}

    def tic():
        torch.cuda.synchronize()
        return time.perf_counter()

    t0 = tic()
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    for idx, br in sc.brdf_groups:
        r.set_wall_brdf(idx, br, sc.directions, sc.directions)
    r.set_air_attenuation(FrequencyData(sc.air_attenuation, sc.frequencies))
    t["host_setup_s"] = tic() - t0
    t0 = tic()
    r.bake_geometry()
    t["bake_s"] = tic() - t0
    t0 = tic()
    r.init_source_energy(Coordinates(*sc.sources[0]))
    t["inject_s"] = tic() - t0
    c, dt_bin = sc.speed_of_sound, sc.time_resolution
    t0 = tic()
    plan = r._exchange_plan(c, dt_bin)
    t["plan_s"] = tic() - t0

    e0d = r._dev["energy_init_source"]
    d0 = r._dev["distance_patches_to_source"]
    P, D, B = e0d.shape
    T = sc.n_samples
    state = ops.ExchangeState(P, D, B, T, dev)
    state.init(e0d, d0, c, dt_bin)
    for _ in range(args.warmup):
        state.step(plan)
    torch.cuda.synchronize()

    lag = plan.lag.to(torch.int64)
    U = int((B * torch.clamp(T - lag, min=0)).sum().item())
    Vd = plan.Vd
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        state.step(plan)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    kern_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_s = float(np.mean(kern_ms)) / 1e3
    state.check()

    # collection after the timed region (reported, not part of value)
    t0 = tic()
    etc = state.finalize(out=state.hi)
    r._etc = etc
    r._speed_of_sound, r._etc_time_resolution = c, dt_bin
    r._etc_duration = sc.duration
    hist = r.collect_energy_receiver_mono(Coordinates(*sc.receivers.T))
    t["collect_s"] = tic() - t0
    assert np.isfinite(hist.time).all()

    value = U * args.steps / wall
    n_state = P * D * B * T
    alg_bytes = 4 * U + Vd * (4 * D * B + 8) + 12 * n_state
    achieved_gbs = alg_bytes / avg_s / 1e9
    flops = 2 * D * U
    line = {
        "metric": "patch-pair·bin updates/s at 5k patches",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 accumulate",
        "data": "synthetic (seeded street-canyon scene)",
        "config": {
            "workload": f"{sc.name}: exchange, 1 source, "
                        f"P={P}, D={D}, B={B}, T={T}, one order per step",
            "patches": P, "directions": D, "bands": B, "bins": T,
            "directed_pairs": Vd, "updates_per_order": U,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "exchange_step_kernel",
            "kernel_ms": avg_s * 1e3,
            "algorithmic_bytes": alg_bytes,
            "valu_tflops": flops / avg_s / 1e12,
        },
        "setup": t,
    }
    if not args.no_cpu_baseline:
        ups, nth, T_s, n_dest = cpu_baseline(plan, D, T, args.cpu_seconds)
        line["cpu_baseline"] = {
            "value": ups, "unit": "updates/s", "cores": nth, "kind": "port",
            "sample": f"FP64 C oracle (oracle/sparrow_oracle.c "
                      f"orc_exchange_order, OpenMP over destinations) on "
                      f"{n_dest} seeded destination patches of the same "
                      f"scene with their full source lists, band 0, "
                      f"{T_s} bins; host: {_cpu_info()}",
        }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
