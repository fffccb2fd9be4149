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
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and dense MFMA), published


def _scene(name, brdf="scattering"):
    from sparrowpy_amd import scenes
    if name == "canyon":
        return scenes.street_canyon(brdf=brdf)
    if name == "shoebox":
        return scenes.directional_shoebox(brdf=brdf)
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


def _generic_rows(plan, D):
    """Host copies of a plan as per-entry rows: (row_ptr, src, lag, dir,
    coef[e, d] for band 0) -- the reference's per-pair update."""
    from sparrowpy_amd import ops
    row_ptr = plan.row_ptr.cpu().numpy()
    lag = plan.lag.cpu().numpy()
    if isinstance(plan, ops.GroupedPlan):
        flag = (lag & ops.GroupedPlan.LAST_IN_GROUP) != 0
        lag = lag & (ops.GroupedPlan.LAST_IN_GROUP - 1)
        rowid = plan.rowid.cpu().numpy().astype(np.int64)
        src, dirs = rowid // D, rowid % D
        gs = plan.grp_start.cpu().numpy()
        gk = plan.grp_key.cpu().numpy()
        g_of_e = np.searchsorted(gs, np.arange(rowid.size), side="right") - 1
        rt = plan.rtab[:, 0, :D].cpu().numpy()
        coef = plan.scal[0].cpu().numpy()[:, None] * rt[gk[g_of_e]]
        if plan.diffuse_specular:
            # entries after a row's last flagged entry have no specular part
            row_of_e = np.repeat(np.arange(plan.P), np.diff(row_ptr))
            after = np.cumsum(flag[::-1])[::-1]
            after_end = np.append(after, 0)[row_ptr[1:]][row_of_e]
            coef[(after - after_end) == 0] = 0.0
            coef += plan.dscal[0].cpu().numpy()[:, None]
    else:
        src = plan.src.cpu().numpy()
        dirs = plan.dir.cpu().numpy()
        coef = plan.coef[:, 0, :D].cpu().numpy()
    return row_ptr, src, lag, dirs, coef


def _host_cores():
    """(threads this process may use, logical CPUs of the host)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or 1


def cpu_baseline(plan, D, T, seconds, nthreads, seed=0):
    """Time the FP64 oracle's exchange order on a bounded sample.

    Sample: a seeded subset of destination patches with their full source
    lists, band 0 only, the histogram truncated to ``T_s`` bins, sized so
    the timed call takes about ``seconds``.  The oracle performs the
    reference's per-pair D-wide update (RadiosityFast.py:1133-1143);
    ``nthreads`` = 1 is the serial loop of the reference's numba.njit
    exchange (RadiosityFast.py:1396), more threads split the destinations.
    Returns (updates/s, T_s, n_dest).
    """
    import oracle as orc
    orc.build()
    lib = orc._load()
    row_ptr, src, lag, dirs, coef_all = _generic_rows(plan, D)
    P = plan.P
    T_s = min(T, 2400)
    rng = np.random.default_rng(seed)
    prev = np.full((P, D, 1, T_s), 1e-3)
    order = rng.permutation(P)

    def run(dests):
        idx = np.concatenate([np.arange(row_ptr[j], row_ptr[j + 1])
                              for j in dests])
        rp = np.zeros(len(dests) + 1, np.int64)
        rp[1:] = np.cumsum([row_ptr[j + 1] - row_ptr[j] for j in dests])
        s_ = np.ascontiguousarray(src[idx], dtype=np.int32)
        l_ = np.ascontiguousarray(lag[idx], dtype=np.int32)
        d_ = np.ascontiguousarray(dirs[idx], dtype=np.int32)
        coef = np.ascontiguousarray(coef_all[idx], dtype=np.float64)
        cur_s = np.zeros((len(dests), D, 1, T_s))
        upd = int(np.sum(np.maximum(0, T_s - l_.astype(np.int64))))
        t0 = time.perf_counter()
        lib.orc_exchange_order(T_s, P, D, 1, prev, cur_s, rp, s_, l_, d_,
                               coef, 0, len(dests), nthreads)
        return upd, time.perf_counter() - t0

    # calibrate on a few destinations, then size the sample to `seconds`
    n0 = max(2, min(8, nthreads))
    upd, dt = run(order[:n0])
    per_dest = max(dt / n0, 1e-6) / max(1, min(nthreads, n0))
    n_dest = int(min(P, max(n0, seconds / per_dest)))
    upd, dt = run(order[:n_dest])
    return upd / dt, T_s, n_dest


def _plan_kind(plan):
    from sparrowpy_amd import ops
    if not isinstance(plan, ops.GroupedPlan):
        return "generic"
    return "ds" if plan.diffuse_specular else "grouped"


def byte_models(plan, P, D, B, T, U, n_pair_per_order):
    """HBM bytes per launch of the order's kernels.

    ``min``: compulsory traffic, what a kernel must move at least: every
    source row (patch, outgoing direction) the plan reads is read once
    (B*T floats each), every output written once, every plan word read
    once.  ``reference``: the accounting of one update per gathered float
    (4 bytes per update, the F~ slice and delay/direction per directed
    pair, the order's histogram written, the total read and written);
    cache reuse lets a kernel beat it, so it is no roofline.
    Returns {kernel: {"min": bytes, "reference": bytes}}, with the
    group-sum round trip of the two-pass plans on its own line.
    """
    import torch
    kind = _plan_kind(plan)
    Vd = plan.Vd
    n_state = P * D * B * T
    total_rw = 20 * n_state * n_pair_per_order  # prev + hi/lo read + write
    out = {}
    if kind == "generic":
        rows = int(torch.unique(plan.src.to(torch.int64) * D +
                                plan.dir.to(torch.int64)).numel())
        meta = Vd * (12 + 8 * D * B) + 8 * (P + 1)
        out["exchange_step_kernel"] = {
            "min": 4 * B * T * rows + meta + 4 * n_state + 16 * n_state,
            "reference": 4 * U + Vd * (4 * D * B + 8) + 12 * n_state}
        return out
    rows = int(torch.unique(plan.rowid).numel())
    G = plan.n_groups
    gsum = 4 * G * B * T
    tiles = getattr(plan, "tiles", None)
    n_w = 2 if kind == "ds" else 1
    if tiles is not None:
        JT, M = tiles.JT, tiles.M
        meta = M * (4 + 4 * JT + 8 * JT * B * n_w) + 8 * (
            tiles.tptr.numel() + n_w * tiles.P)
        gname = "tile_lds_gather_kernel" if JT == 16 else \
            "tile_gather_kernel"
    else:
        meta = Vd * (8 + 8 * B * n_w) + 8 * (2 * P + 2)
        gname = "ds_gather_kernel" if kind == "ds" else "group_gather_kernel"
    out[gname] = {"min": 4 * B * T * rows + meta + gsum,
                  "reference": 4 * U + Vd * (8 + 8 * B * n_w) + gsum,
                  "group_sums": gsum}
    sname = "ds_spread_kernel" if kind == "ds" else "group_spread_kernel"
    spread_meta = 4 * G + 8 * (P + 1) + (
        0 if kind == "ds" else plan.rtab.numel() * 8)
    out[sname] = {"min": gsum + 4 * n_state + total_rw + spread_meta,
                  "reference": gsum + 4 * n_state + total_rw,
                  "group_sums": gsum}
    return out


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                               "profiles", "traffic_street_canyon.json")
KERNEL_SOURCES = ["exchange.hip", "tiles.hip", "spr_exchange.hpp",
                  "spr_util.hpp", "spr_common.hpp"]


def _code_only(text):
    """C++ source without comments and with whitespace runs collapsed: the
    part of a kernel source that decides what the compiler builds."""
    import re
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return " ".join(text.split())


def kernel_sources_sha():
    """sha256 over the exchange kernels' sources (their code, comments
    left out): a counter profile is only used for the kernels it was
    measured on."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "sparrowpy_amd", "csrc")
    for name in KERNEL_SOURCES:
        with open(os.path.join(csrc, name)) as f:
            h.update(name.encode() + b"\0" + _code_only(f.read()).encode())
    return h.hexdigest()[:16]


def measured_traffic(kernel, config):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE passes,
    scaled by a copy of known size), or (None, reason) when the profile
    covers another workload or other kernel sources."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, "no counter profile"
    if any(prof["workload"].get(k) != config.get(k)
           for k in prof["workload"]):
        return None, "counter profile of another workload"
    if prof.get("kernel_sources_sha") != kernel_sources_sha():
        return None, ("counter profile measured on other kernel sources "
                      "(stale; regenerate with bench.py --traffic-from)")
    for name, v in prof["kernels"].items():
        if name.split("<")[0] == kernel:
            return v["bytes"], None
    return None, f"{kernel} not in the counter profile"


def adjoint_timing(log):
    """Forward (with order checkpoints) and adjoint time of the receiver
    histograms' gradient with respect to the BRDF tables, on the 10x8x3 m
    directional shoebox (1,072 patches, one BRDF per patch, 64 directions,
    6 bands, 6,000 bins, order 30, checkpoints every 5 orders, 4
    receivers, log-energy loss against perturbed tables)."""
    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, adjoint
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData
    sc = _scene("shoebox")
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    for idx, br in sc.brdf_groups:
        r.set_wall_brdf(idx, br, sc.directions, sc.directions)
    r.set_air_attenuation(FrequencyData(sc.air_attenuation, sc.frequencies))
    r.bake_geometry()
    adj = adjoint.ExchangeAdjoint(
        r, Coordinates(*sc.sources[0]), Coordinates(*sc.receivers.T),
        sc.speed_of_sound, sc.time_resolution, sc.duration, sc.max_order,
        checkpoint_every=5)
    g = torch.Generator(device="cuda").manual_seed(0)
    A0, C0 = adj.A0, adj.C0
    target = adj.forward(A0 * (1 + 0.1 * torch.rand(A0.shape, generator=g,
                                                      device="cuda",
                                                      dtype=A0.dtype)),
                         C0).to(torch.float64)
    adj.release()

    def once():
        A = A0.clone().requires_grad_(True)
        C = C0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        H = adjoint.receiver_histograms(adj, A, C)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        adjoint.log_energy_loss(H, target).backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, float(A.grad.abs().max())

    once()  # warm-up
    fwd, bwd, gmax = once()
    log(f"adjoint: forward {fwd:.3f} s, backward {bwd:.3f} s")
    out = {"workload": f"{sc.name}: P={adj.P}, D={adj.D}, B={adj.B}, "
                       f"T={adj.T}, order {adj.K}, checkpoints every 5, "
                       f"{adj.R} receivers",
           "forward_s": fwd, "adjoint_s": bwd, "adjoint_over_forward":
           bwd / fwd, "max_abs_grad": gmax,
           "plan": "diffuse+specular", "directed_pairs": adj.plan.Vd}
    del adj, r
    torch.cuda.empty_cache()
    return out


def adjoint_brdf_timing(log):
    """The same for BRDF rows of no diffuse + specular form: the 10x8x3 m
    shoebox with one directional-scattering BRDF per wall (the same 1,072
    patches, source and receivers), gradient with respect to the whole
    BRDF table on the grouped plan (BRDFExchangeAdjoint)."""
    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, adjoint, scenes
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData
    sc = scenes.directional_shoebox(brdf="directional")
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    for idx, br in sc.brdf_groups:
        r.set_wall_brdf(idx, br, sc.directions, sc.directions)
    r.set_air_attenuation(FrequencyData(sc.air_attenuation, sc.frequencies))
    r.bake_geometry()
    adj = adjoint.BRDFExchangeAdjoint(
        r, Coordinates(*sc.sources[0]), Coordinates(*sc.receivers.T),
        sc.speed_of_sound, sc.time_resolution, sc.duration, sc.max_order,
        checkpoint_every=5)
    g = torch.Generator(device="cuda").manual_seed(0)
    t0 = adj.table0
    target = adj.forward(t0 * (1 + 0.1 * torch.rand(
        t0.shape, generator=g, device="cuda", dtype=t0.dtype))).to(
            torch.float64)
    adj.release()

    def once():
        table = t0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        a = time.perf_counter()
        H = adjoint.receiver_histograms_brdf(adj, table)
        torch.cuda.synchronize()
        b = time.perf_counter()
        adjoint.log_energy_loss(H, target).backward()
        torch.cuda.synchronize()
        c = time.perf_counter()
        return b - a, c - b, float(table.grad.abs().max())

    once()  # warm-up
    fwd, bwd, gmax = once()
    log(f"adjoint (general BRDF rows): forward {fwd:.3f} s, backward "
        f"{bwd:.3f} s")
    out = {"workload": f"{sc.name}: P={adj.P}, D={adj.D}, B={adj.B}, "
                       f"T={adj.T}, order {adj.K}, checkpoints every 5, "
                       f"{adj.R} receivers, {adj.NK} BRDF rows",
           "forward_s": fwd, "adjoint_s": bwd, "adjoint_over_forward":
           bwd / fwd, "max_abs_grad": gmax, "plan": "grouped",
           "directed_pairs": adj.plan.Vd, "groups": adj.plan.n_groups}
    del adj, r
    torch.cuda.empty_cache()
    return out

def bench_inverse(args, log):
    """Inverse materials on the directional shoebox: target histograms from
    the scene's seeded per-patch absorption and scattering, then Adam steps
    (``--steps``, default 5; 200 for the full study) through the adjoint
    from perturbed materials.  A step = forward with checkpoints + adjoint +
    Adam update of 2 x 1,072 x 6 parameters."""
    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, adjoint
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData
    sc = _scene("shoebox")
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    for idx, br in sc.brdf_groups:
        r.set_wall_brdf(idx, br, sc.directions, sc.directions)
    r.set_air_attenuation(FrequencyData(sc.air_attenuation, sc.frequencies))
    r.bake_geometry()
    adj = adjoint.ExchangeAdjoint(
        r, Coordinates(*sc.sources[0]), Coordinates(*sc.receivers.T),
        sc.speed_of_sound, sc.time_resolution, sc.duration, sc.max_order,
        checkpoint_every=5)
    ta, ts = sc.materials
    sv, _ = adjoint.specular_factors(sc.directions)
    kw = dict(dtype=torch.float64, device="cuda")
    A, C = adjoint.scattering_tables(torch.tensor(ta, **kw),
                                     torch.tensor(ts, **kw), sv)
    target = adj.forward(A, C).to(torch.float64)
    adj.release()
    # start: every patch's absorption 0.05 too high and scattering 0.1 too
    # low, plus per-patch noise (the per-band means are what 4 receivers
    # identify; the per-patch values are not determined by them)
    rng = np.random.default_rng(11)
    a0 = np.clip(ta + .05 + rng.normal(0, .03, ta.shape), .01, .99)
    s0 = np.clip(ts - .1 + rng.normal(0, .05, ts.shape), .01, .99)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a, s, losses = adjoint.recover_materials(adj, target, a0, s0, sv,
                                             steps=args.steps, lr=0.01,
                                             log=log)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    a, s = a.cpu().numpy(), s.cpu().numpy()
    return {
        "metric": "inverse-materials Adam steps/s (forward + adjoint)",
        "value": args.steps / wall, "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 0,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 state, f64 accumulate",
        "data": "synthetic (seeded directional shoebox)",
        "config": {"workload": f"{sc.name}: P={adj.P}, D={adj.D}, "
                               f"B={adj.B}, T={adj.T}, order {adj.K}, "
                               f"checkpoints every 5, {adj.R} receivers, "
                               f"{a.size + s.size} parameters"},
        "loss_first": losses[0], "loss_last": losses[-1],
        "loss_every_20": losses[::20],
        "abs_err_alpha": [float(np.abs(a0 - ta).mean()),
                          float(np.abs(a - ta).mean())],
        "abs_err_scattering": [float(np.abs(s0 - ts).mean()),
                               float(np.abs(s - ts).mean())],
        "band_mean_err_alpha": [np.abs(a0.mean(0) - ta.mean(0)).tolist(),
                                np.abs(a.mean(0) - ta.mean(0)).tolist()],
        "band_mean_err_scattering": [
            np.abs(s0.mean(0) - ts.mean(0)).tolist(),
            np.abs(s.mean(0) - ts.mean(0)).tolist()],
    }


def bench_dense(args, log):
    """The reference-shaped entry point spr_energy_exchange (dense F~
    [P,P,D,B] in float64, RadiosityFast.py:1073-1145) on the directional
    shoebox with per-wall directional-scattering BRDFs: time per order of
    the grouped exchange it runs on factorising F~ rows, and of the
    per-pair plan it falls back to, timed on the same F~ with every row
    perturbed by up to 1e-3 (no longer factorising).  Per order = (time of
    3 orders - time of 1 order) / 2, each call including its plan."""
    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, ops
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData
    sc = _scene("shoebox", "directional")
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    for idx, br in sc.brdf_groups:
        r.set_wall_brdf(idx, br, sc.directions, sc.directions)
    r.set_air_attenuation(FrequencyData(sc.air_attenuation, sc.frequencies))
    r.bake_geometry()
    r.init_source_energy(Coordinates(*sc.sources[0]))
    b, g = r._bake, r._geometry()
    P, D, B, T = r.n_patches, int(b["D"]), int(b["n_bins"]), sc.n_samples
    c, dt = sc.speed_of_sound, sc.time_resolution
    fft = ops.form_factors_with_directivity_dim(
        b["vis"], b["ff"], b["n_bins"], g["centers"], g["areas"], b["air"],
        g["wall_ids"], b["brdf"], b["brdf_index"], b["sources"],
        b["receivers"])
    ctr = r.patches_center
    dd = ctr[:, None, :] - ctr[None, :, :]
    dist = np.sqrt((dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1])
                   + dd[..., 2] * dd[..., 2])
    dist = torch.tensor(dist, dtype=torch.float64, device="cuda")
    e0d = r._dev["energy_init_source"]
    d0 = r._dev["distance_patches_to_source"]
    V = int(b["pairs"].shape[0])

    def run(K):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = ops.energy_exchange(T, e0d, d0, dist, fft, b["p2o"], c, dt, K,
                                  b["pairs"])
        torch.cuda.synchronize()
        del out
        return time.perf_counter() - t0, ops.energy_exchange_groups()

    res = {}
    for kind in ("grouped", "per_pair"):
        if kind == "per_pair":
            gen = torch.Generator(device="cuda").manual_seed(0)
            fft.mul_(1 + 1e-3 * torch.rand(fft.shape, generator=gen,
                                           device="cuda", dtype=fft.dtype))
        run(1)  # warm-up
        t1, groups = run(1)
        t3, _ = run(3)
        per = (t3 - t1) / 2
        res[kind] = {"ms_per_order": per * 1e3,
                     "plan_ms": (t1 - per) * 1e3, "groups": groups}
        log(f"dense entry, {kind}: {per * 1e3:.1f} ms per order, plan "
            f"{(t1 - per) * 1e3:.1f} ms, {groups} groups")
    return {
        "metric": "dense-F~ exchange time per order (spr_energy_exchange)",
        "value": res["grouped"]["ms_per_order"], "unit": "ms/order",
        "n_gpus": 1, "higher_is_better": False, "vs_baseline": None,
        "dtype": "f32 state, f64 accumulate",
        "data": "synthetic (seeded directional shoebox)",
        "config": {"workload": f"{sc.name}: P={P}, D={D}, B={B}, T={T}, "
                               f"{2 * V} directed pairs, dense F~ "
                               f"{P * P * D * B * 8 / 2**30:.1f} GiB"},
        "grouped": res["grouped"], "per_pair": res["per_pair"],
        "speedup": res["per_pair"]["ms_per_order"] /
        res["grouped"]["ms_per_order"],
    }


def _tiling_str(t):
    rows, block, r = t
    if not rows:
        return "row by row"
    blk = "consecutive" if block is None else f"{block[0]}x{block[1]} blocks"
    return f"{rows} rows, {blk}, {64 * r}-bin blocks"


def bench_batched(args, log):
    """Many rooms and sources: the per-instance loop of the reference API
    against the batched, graph-captured exchange on the same baked scenes.
    Throughput = exchange updates of all instances / wall time of the
    exchange and receiver collection (bake excluded, reported)."""
    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, scenes
    from sparrowpy_amd.batch import BatchedExchange, BatchItem, \
        run_per_instance
    from sparrowpy_amd.coordinates import FrequencyData
    scs = scenes.batched_shoeboxes(args.rooms)
    t0 = time.perf_counter()
    items = []
    for sc in scs:
        r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
        for idx, br in sc.brdf_groups:
            r.set_wall_brdf(idx, br, sc.directions, sc.directions)
        r.set_air_attenuation(FrequencyData(sc.air_attenuation,
                                            sc.frequencies))
        r.bake_geometry()
        items.append(BatchItem(r, sc.sources, sc.receivers))
    torch.cuda.synchronize()
    bake_s = time.perf_counter() - t0
    n_inst = sum(len(sc.sources) for sc in scs)
    log(f"baked {len(scs)} rooms ({n_inst} instances) in {bake_s:.1f} s")
    c, dt, dur, K = (scs[0].speed_of_sound, scs[0].time_resolution,
                     scs[0].duration, scs[0].max_order)
    # the batched run first: its chunk buffers are tens of GiB each, and
    # the per-instance loop leaves the device heap fragmented (each room's
    # plan is placed in the previous room's freed histograms)
    batched_s = loop_s = float("nan")
    out = ref = None
    if args.batched_part in ("both", "batched"):
        be = BatchedExchange(
            items, c, dt, dur, K, tile_rows=args.batch_tile_rows,
            tile_block=(None if args.batch_tile_block == "none" else
                        "auto" if args.batch_tile_block == "auto" else tuple(
                            int(v) for v in args.batch_tile_block.split("x"))),
            tile_r=args.tile_r if args.tile_r is not None else 8)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = be.run()
        torch.cuda.synchronize()
        batched_s = time.perf_counter() - t0
        log(f"batched {batched_s:.2f} s, classes {be.choice}, tiling "
            f"{_tiling_str(be.tiling)} of {len(be.tiling_times)} timed")
        be.pool = None
        torch.cuda.empty_cache()
    if args.batched_part in ("both", "loop"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref = run_per_instance(items, c, dt, dur, K)
        torch.cuda.synchronize()
        loop_s = time.perf_counter() - t0
        log(f"per-instance loop {loop_s:.2f} s")
    if out is None or ref is None:
        U = be.updates if out is not None else None
        return {"metric": "batched rooms, one part only",
                "part": args.batched_part, "batched_s": batched_s,
                "loop_s": loop_s, "updates": U, "n_gpus": 1,
                "instances": n_inst}
    err = 0.0
    for key, v in ref.items():
        a, b = out[key].double(), v.double()
        peak = b.abs().amax(dim=-1, keepdim=True)
        mask = b.abs() >= peak * 1e-6
        if mask.any():
            err = max(err, float(((a - b).abs() / b.abs())[mask].max()))
    U = be.updates
    return {
        "metric": "patch-pair·bin updates/s, batched rooms",
        "value": U / batched_s, "unit": "updates/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": batched_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 state, f64 accumulate",
        "data": "synthetic (seeded shoeboxes)",
        "config": {"workload": f"{len(scs)} shoeboxes [3,20]x[3,20]x"
                               f"[2.5,6] m, patch [0.5,1.5] m, 1-4 sources,"
                               f" 1-8 receivers, D 16/64, 1-8 bands, "
                               f"2 kHz x 1 s, order {K}",
                   "rooms": len(scs), "instances": n_inst,
                   "patches": int(sum(it.radiosity.n_patches
                                      for it in items)),
                   "updates": U,
                   "plan_per_class": {f"D{k[0]}xB{k[1]}": v
                                      for k, v in be.choice.items()},
                   "tiling": _tiling_str(be.tiling),
                   "tiling_candidates_ms": {
                       _tiling_str(t): v * 1e3
                       for t, v in be.tiling_times.items()}},
        "per_instance_loop": {"s": loop_s, "updates/s": U / loop_s},
        "speedup_vs_loop": loop_s / batched_s,
        "max_rel_err_vs_loop": err, "bake_s": bake_s,
        "includes": "exchange, order-0 injection, receiver collection; "
                    "batched also plan concatenation, autotuning and "
                    "graph capture",
    }


def _kernel_name(full):
    """Kernel name with its template arguments, without return type,
    namespaces or parameter list, as rocprofv3 prints it demangled:
    "void ns::(anonymous namespace)::k<8, 16>(long, ...)" -> "k<8, 16>"."""
    s = full.replace("(anonymous namespace)::", "").split("(")[0].strip()
    if s.startswith("void "):
        s = s[len("void "):]
    head, sep, tail = s.partition("<")
    return head.split("::")[-1] + sep + tail


def _counter_means(root, counter):
    """{kernel name: mean value of ``counter`` per dispatch} from the
    rocprofv3 CSV output under ``root`` (``*counter_collection.csv``)."""
    import csv
    import glob
    vals = {}
    files = glob.glob(os.path.join(root, "**", "*counter_collection.csv"),
                      recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {root}")
    for path in files:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name = _kernel_name(row["Kernel_Name"])
                key = (name, row.get("Dispatch_Id", len(vals)))
                vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
    per = {}
    for (name, _), v in vals.items():
        per.setdefault(name, []).append(v)
    return {k: float(np.mean(v)) for k, v in per.items()}


def traffic_from(dirs, workload):
    """Write the counter profile from four rocprofv3 runs: FETCH_SIZE and
    WRITE_SIZE passes of the benchmark and of ``--copy-calibration`` (a
    device copy of known size, which fixes the counters' scale)."""
    fetch_dir, write_dir, copyf_dir, copyw_dir = dirs
    copy_name = "__amd_rocclr_copyBuffer"
    cf = _counter_means(copyf_dir, "FETCH_SIZE")
    cw = _counter_means(copyw_dir, "WRITE_SIZE")
    n = 1 << 32
    fetch_scale = n / (cf[copy_name] * 1024)
    write_scale = n / (cw[copy_name] * 1024)
    fe = _counter_means(fetch_dir, "FETCH_SIZE")
    wr = _counter_means(write_dir, "WRITE_SIZE")
    kernels = {}
    for name in sorted(set(fe) & set(wr)):
        fb = fe[name] * 1024 * fetch_scale
        wb = wr[name] * 1024 * write_scale
        kernels[name] = {"fetch_bytes": fb, "write_bytes": wb,
                         "bytes": fb + wb, "fetch_kib_raw": fe[name],
                         "write_kib_raw": wr[name]}
    out = {"workload": workload, "kernel_sources_sha": kernel_sources_sha(),
           "fetch_scale": fetch_scale, "write_scale": write_scale,
           "kernels": kernels}
    with open(TRAFFIC_PROFILE, "w") as f:
        json.dump(out, f, indent=1)
    return out


def copy_calibration(n_bytes=1 << 32, reps=3):
    """Copy n_bytes device-to-device ``reps`` times (one kernel each)."""
    import torch
    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n_bytes, "copy_bytes_written":
                      n_bytes, "reps": reps}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="canyon",
                    choices=["canyon", "shoebox", "small", "batched",
                             "inverse", "adjoint", "dense"],
                    help="adjoint: only the adjoint/forward timings of the "
                         "default run (diffuse + specular and general "
                         "BRDF rows); dense: the reference-shaped "
                         "spr_energy_exchange on a dense F~")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rooms", type=int, default=256,
                    help="rooms of --config batched")
    ap.add_argument("--batched-part", default="both",
                    choices=["both", "batched", "loop"],
                    help="--config batched: run both paths, or only one "
                         "(to profile it alone)")
    ap.add_argument("--batch-tile-rows", type=int, default=16,
                    choices=[0, 4, 8, 16],
                    help="--config batched: destination rows per tile of "
                         "the batched classes' gathers")
    ap.add_argument("--batch-tile-block", default="auto",
                    help="--config batched: tiles as AxB blocks of each "
                         "room's patches (e.g. 4x4 with --batch-tile-rows "
                         "16), none, or auto (the class default)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward timing")
    ap.add_argument("--brdf", default="scattering",
                    choices=["scattering", "directional"],
                    help="--config canyon: the walls' BRDFs from scattering "
                         "coefficients (constant + specular rows) or from "
                         "seeded directional scattering coefficients "
                         "(general rows, shared per wall)")
    ap.add_argument("--plan", default="auto",
                    choices=["auto", "ds", "grouped", "generic"],
                    help="exchange plan (auto: grouped when sources share "
                         "BRDF rows)")
    ap.add_argument("--tile-rows", type=int, default=None, choices=[0, 4, 8, 16],
                    help="destination rows per tile of the grouped gather "
                         "(default: the class default; 0: row by row)")
    ap.add_argument("--tile-block", default=None,
                    help="tiles as AxB blocks of neighbouring patches "
                         "(e.g. 4x4 with --tile-rows 16); default: the "
                         "class default")
    ap.add_argument("--tile-r", type=int, default=None,
                    choices=[16, 8],
                    help="16-row tiles: bins per lane of the gather")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a device copy of known size (to scale "
                         "the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--traffic-from", nargs=4, default=None,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR",
                             "COPY_WRITE_DIR"),
                    help="write profiles/traffic_street_canyon.json from "
                         "rocprofv3 counter runs of this benchmark and of "
                         "--copy-calibration, then exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        return copy_calibration()
    if args.traffic_from:
        from sparrowpy_amd import geometry
        sc = _scene(args.config, args.brdf)
        wp = np.array([w.pts for w in sc.walls])
        wn = np.array([w.normal for w in sc.walls])
        n_patches = int(geometry.process_patches(wp, wn,
                                                 sc.patch_size)[2])
        workload = {"patches": n_patches, "directions":
                    int(sc.directions.cartesian.shape[0]),
                    "bands": len(sc.frequencies), "bins": sc.n_samples}
        prof = traffic_from(args.traffic_from, workload)
        print(json.dumps({k: v["bytes"] for k, v in
                          prof["kernels"].items()}, indent=1))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not "
              "supported (the exchange shards by band with no collectives; "
              "only the one-GPU measurement is implemented)", file=sys.stderr)
        return 2

    import torch
    from sparrowpy_amd import DirectionalRadiosityFast, ops
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData

    if args.config in ("batched", "inverse", "dense"):
        def blog(msg):
            print(f"[bench] {msg}", file=sys.stderr, flush=True)
        fn = {"batched": bench_batched, "inverse": bench_inverse,
              "dense": bench_dense}[args.config]
        print(json.dumps(fn(args, blog)), flush=True)
        return 0
    if args.config == "adjoint":
        # the two adjoint timings of the default run, alone (to profile them)
        def alog(msg):
            print(f"[bench] {msg}", file=sys.stderr, flush=True)
        ds, gen = adjoint_timing(alog), adjoint_brdf_timing(alog)
        print(json.dumps({
            "metric": "adjoint / forward time", "value": gen["adjoint_s"] /
            gen["forward_s"], "unit": "ratio", "n_gpus": 1,
            "higher_is_better": False, "data": "synthetic (seeded "
            "directional shoebox)", "adjoint": ds, "adjoint_brdf": gen}),
            flush=True)
        return 0
    sc = _scene(args.config, args.brdf)
    torch.cuda.init()
    dev = torch.device("cuda", 0)
    t = {}

    def tic():
        torch.cuda.synchronize()
        return time.perf_counter()

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = tic()
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    for idx, br in sc.brdf_groups:
        r.set_wall_brdf(idx, br, sc.directions, sc.directions)
    r.set_air_attenuation(FrequencyData(sc.air_attenuation, sc.frequencies))
    if args.tile_rows is not None:
        r.TILE_ROWS = args.tile_rows
    if args.tile_r is not None:
        r.TILE_R = args.tile_r
    if args.tile_block is not None:
        blk = args.tile_block.rstrip("u")
        r.TILE_BLOCK = (None if blk == "none" else
                        tuple(int(x) for x in blk.split("x")))
        # a trailing "u": blocks run together without padding
        r.TILE_PAD = not args.tile_block.endswith("u")
    t["host_setup_s"] = tic() - t0
    t0 = tic()
    r.bake_geometry()
    t["bake_s"] = tic() - t0
    log(f"bake {t['bake_s']:.2f} s, P={r.n_patches}")
    t0 = tic()
    r.init_source_energy(Coordinates(*sc.sources[0]))
    t["inject_s"] = tic() - t0
    c, dt_bin = sc.speed_of_sound, sc.time_resolution
    t0 = tic()
    plan = r._exchange_plan(c, dt_bin,
                            None if args.plan == "auto" else args.plan)
    kind = ("generic" if not isinstance(plan, ops.GroupedPlan) else
            "ds" if plan.diffuse_specular else "grouped")
    t["plan_s"] = tic() - t0
    tiles = getattr(plan, "tiles", None)
    log(f"plan {t['plan_s']:.2f} s ({kind}), directed pairs={plan.Vd}"
        + (f", tiles of {tiles.JT} rows: {tiles.M} merged entries"
           if tiles is not None else ""))

    e0d = r._dev["energy_init_source"]
    d0 = r._dev["distance_patches_to_source"]
    P, D, B = e0d.shape
    T = sc.n_samples
    state = ops.ExchangeState(P, D, B, T, dev)
    state.init(e0d, d0, c, dt_bin)
    for _ in range(args.warmup):
        state.step(plan)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} orders done")

    lag = plan.lag.to(torch.int64) & (ops.GroupedPlan.LAST_IN_GROUP - 1)
    U = int((B * torch.clamp(T - lag, min=0)).sum().item())
    Vd = plan.Vd
    stream = torch.cuda.current_stream()
    two_pass = kind in ("ds", "grouped")
    n_ev = 3 if two_pass else 2
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n_ev)]
          for _ in range(args.steps)]
    n_pair = 0
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        if two_pass:
            state.gather(plan)
            ev[k][1].record(stream)
            n_pair += state.pending     # this spread adds two orders
            state.spread(plan)
            state.prev, state.cur = state.cur, state.prev
        else:
            state.step(plan)
        ev[k][-1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    log(f"timed {args.steps} orders: {wall:.3f} s")
    step_s = float(np.mean([e[0].elapsed_time(e[-1]) for e in ev])) / 1e3
    if two_pass:
        gather_s = float(np.mean([e[0].elapsed_time(e[1]) for e in ev])) / 1e3
        spread_s = float(np.mean([e[1].elapsed_time(e[2]) for e in ev])) / 1e3
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
    models = byte_models(plan, P, D, B, T, U, n_pair / args.steps)
    kernels = {}
    if two_pass:
        gather_name, spread_name = list(models)
        kernels = {
            gather_name: dict(ms=gather_s * 1e3, **models[gather_name]),
            spread_name: dict(ms=spread_s * 1e3, **models[spread_name]),
        }
        dom, dom_s = gather_name, gather_s
    else:
        dom, dom_s = "exchange_step_kernel", step_s
        kernels = {dom: dict(ms=step_s * 1e3, **models[dom])}
    for k in kernels.values():
        k["GB/s_min"] = k["min"] / (k["ms"] / 1e3) / 1e9
        k["GB/s_reference"] = k["reference"] / (k["ms"] / 1e3) / 1e9
    dom_min = models[dom]["min"]
    achieved_gbs = dom_min / dom_s / 1e9
    # the whole order: compulsory bytes without the group sums (which a
    # one-pass exchange would not move), the group-sum round trip apart
    gsum_rt = 2 * models[dom].get("group_sums", 0)
    order_min = sum(m["min"] for m in models.values()) - gsum_rt
    order_ref = 4 * U + Vd * (4 * D * B + 8) + 12 * n_state
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
            "plan": kind,
            "groups": plan.n_groups if two_pass else None,
            "tile_rows": tiles.JT if tiles is not None else 0,
            "merged_entries": tiles.M if tiles is not None else None,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": None,
            "traffic_source": None,
            "kernel": dom, "kernel_ms": dom_s * 1e3,
            "algorithmic_bytes": dom_min,
            "bytes_model": "compulsory: each source row the plan reads "
                           "once (B*T floats), each output written once, "
                           "plan metadata once",
            "reference_equivalent_bytes": models[dom]["reference"],
            "reference_equivalent_GB/s":
                models[dom]["reference"] / dom_s / 1e9,
        },
        "order": {
            "ms": step_s * 1e3,
            "min_bytes": order_min,
            "group_sum_round_trip_bytes": gsum_rt,
            "GB/s_min": order_min / step_s / 1e9,
            "frac_hbm_min": order_min / step_s / 1e9 / HBM_PEAK_GBS,
            "frac_hbm_min_with_group_sums":
                (order_min + gsum_rt) / step_s / 1e9 / HBM_PEAK_GBS,
            "reference_equivalent_bytes": order_ref,
            "reference_equivalent_GB/s": order_ref / step_s / 1e9,
            "reference_flops_tflops": flops / step_s / 1e12,
            "kernels": kernels,
        },
        "setup": t,
    }
    if dom in ("tile_lds_gather_kernel", "tile_gather_kernel",
               "group_gather_kernel"):
        # the gather of the grouped plan: one FP64 FMA per (entry, band,
        # bin), i.e. per update, plus one f32 -> f64 conversion of the
        # source bin on the same vector ALU
        f64 = 2.0 * U
        line["roofline"]["fp64_vector"] = {
            "flops": f64, "achieved_tflops": f64 / dom_s / 1e12,
            "peak_tflops": FP64_PEAK_TFLOPS,
            "frac": f64 / dom_s / 1e12 / FP64_PEAK_TFLOPS,
            "model": "one FP64 FMA per update (each with an f32->f64 "
                     "conversion on the same VALU)"}
    traffic, why = measured_traffic(dom, line["config"])
    if traffic is not None:
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_source"] = (
            "profiles/traffic_street_canyon.json (rocprofv3 FETCH_SIZE and "
            "WRITE_SIZE passes scaled by a copy of known size, per launch)")
        line["roofline"]["traffic_GB/s"] = traffic / dom_s / 1e9
    else:
        line["roofline"]["traffic_source"] = why
    if not args.no_adjoint:
        del state, etc, hist, r
        torch.cuda.empty_cache()
        line["adjoint"] = adjoint_timing(log)
        line["adjoint_brdf"] = adjoint_brdf_timing(log)
    if not args.no_cpu_baseline:
        usable, logical = _host_cores()
        nth = int(os.environ.get("OMP_NUM_THREADS", usable))
        log(f"cpu baseline sample, {nth} threads")
        ups, T_s, n_dest = cpu_baseline(plan, D, T, args.cpu_seconds, nth)
        log("cpu baseline sample, serial")
        ups1, T_s1, n_dest1 = cpu_baseline(plan, D, T,
                                           args.cpu_seconds / 2, 1, seed=1)
        host = (f"{_cpu_info()}; {logical} logical CPUs on the host, "
                f"{usable} usable by this process")
        line["cpu_baseline"] = {
            "value": ups, "unit": "updates/s", "cores": nth, "kind": "port",
            "sample": f"FP64 C oracle (oracle/sparrow_oracle.c "
                      f"orc_exchange_order: the reference's per-pair D-wide "
                      f"update, OpenMP over destinations, {nth} threads) on "
                      f"{n_dest} seeded destination patches of the same "
                      f"scene with their full source lists, band 0, "
                      f"{T_s} bins; host: {host}",
            "serial": {
                "value": ups1, "unit": "updates/s", "cores": 1,
                "sample": f"the same loop on one thread (the reference's "
                          f"numba.njit exchange is serial, "
                          f"RadiosityFast.py:1396) on {n_dest1} "
                          f"destinations, band 0, {T_s1} bins"},
            "host_logical_cpus": logical, "usable_cpus": usable,
        }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
