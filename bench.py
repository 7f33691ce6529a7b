#!/usr/bin/env python3
"""Throughput of the lattice update sweep on one MI355X.

A step is one compound sweep of the 96^3 SU(2)xU(1)+Higgs lattice
(betaG=9, betaY=27, r=1, msq=-0.20, lambda=0.015, cold start phi=(1,0,0,0)):
1 link heatbath + 4 link overrelaxation rounds + U(1) Metropolis +
4 Higgs overrelaxation + 1 Higgs Metropolis, i.e. 46 checkerboard passes and
23 V variable updates.  `value` is FP32-mixed site-updates/s (FP32 fields and
staples, FP64 reductions), timed with the fields resident in HBM; the same
line reports the plain FP64 rate, the launch-by-launch rate, a batched scan
(96 lattices of 16^3..48^3 in one launch per pass), per-kernel HIP-event
times, the HBM roofline of the dominant kernel and the reference's CPU
build timed on this host.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20]
"""
import argparse
import ctypes as C
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CPU_SHARE = 16  # host cores that go with one GPU of the box


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def slot_params(b, slot=0):
    """The parameters a handle resolved for a slot (link storage included)."""
    from su2higgs_amd import abi
    p = abi.Params()
    abi.check(b.lib.su2hip_slot_params(b.h, slot, C.byref(p)))
    return p


def workload_params(precision, L=96):
    from su2higgs_amd.abi import FP32, make_params
    return make_params(L=(L, L, L), nhiggs=1, u1=1, precision=precision, betasu2=9.0,
                       betau1=27.0, r_u1=1.0, msq_phi=-0.20, lambda_phi=0.015, link_overrelax=4,
                       scalar_sweeps=4, update_su2doublet=1, update_links=1,
                       reunitarize=1 if precision == FP32 else 0, seed=20240096)


def passes_and_updates(p):
    """(passes per sweep, variable updates per sweep / V)."""
    links = p.update_links * (1 + p.link_overrelax + (1 if p.u1 else 0)) * 6
    scal = 0
    if p.nhiggs:
        scal += (p.scalar_sweeps + (1 if p.algorithm_su2doublet == 3 else 0)) * p.update_su2doublet * 2
    return links + scal, (links + scal) / 2


def algo_words(kind, p, hop=True):
    """Compulsory HBM words per pass / V (each datum read or written once),
    for the handle's resolved parameters p (slot_params): a link is 4 words,
    3 in the packed FP32 form (three components, the fourth rebuilt from
    the unit norm); hop: the handle keeps the hopping-term cache
    (Batch.hop_cache())."""
    from su2higgs_amd import abi
    nh, u1, tr, sg = p.nhiggs, p.u1, p.triplet, p.singlet
    packed = p.precision == abi.FP32 and p.link_format == abi.LINKS_PACKED
    lw = 0.75 if packed else 1.0  # link words / 4
    if kind == "hop":
        # the hopping-term cache of one doublet (k_hop): phi 4, alpha 3, the
        # write 12
        return 4 + 3 + 12 if nh == 1 and hop else None
    if kind in (abi.K_LINK_HEATBATH, abi.K_LINK_METROPOLIS, abi.K_LINK_OVERRELAX):
        # U_mu(1-p) 2, U_nu 8, the write 2; with one doublet its cached
        # hopping term H_mu(p) 2, else phi 4 and alpha_mu(p) 0.5; Sigma
        w = (2 + 8 + 2) * lw + (2 if nh == 1 and hop else 4 * nh + 0.5 * u1) + 3 * tr
        if kind != abi.K_LINK_HEATBATH:
            w += 2 * lw  # the old link U_mu(p)
        return w
    if kind == abi.K_U1_METROPOLIS:
        return 3 + 4 * nh + 2 * lw + 0.5
    if kind in (abi.K_DOUBLET_OVERRELAX, abi.K_DOUBLET_METROPOLIS):
        return 4 + 12 * lw + 2 + 3 * u1 + 1.5 * tr + 0.5 * sg
    if kind in (abi.K_TRIPLET_OVERRELAX, abi.K_TRIPLET_METROPOLIS):
        return 3 + 12 * lw + 1.5 + 2 * nh
    if kind in (abi.K_SINGLET_OVERRELAX, abi.K_SINGLET_METROPOLIS):
        return 1 + 0.5 + 2 * nh
    return None


def load_traffic(kernel, workload, precision, packed):
    """HBM bytes per launch of one pass kind, measured by rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, calibrated) and summarised in
    profiles/hbm_traffic.json: the row of that workload, precision and link
    form with the most launches (the chunk the profiled run kept)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rows = json.load(f)["kernels"]
        cand = [r for r in rows if r["workload"] == workload and r["kernel"] == kernel
                and r["precision"] == precision and r["hbm_MB"] is not None
                and bool((r["model_bits"] or 0) & 64) == packed]
        if not cand:
            return None
        return float(max(cand, key=lambda r: r["launches"])["hbm_MB"]) * 1e6
    except (OSError, ValueError, KeyError, TypeError):
        return None


OBS_L = 48          # extent of the observable chains (config-2 couplings and sweep mix)
OBS_CHAINS = 32     # independent chains per precision, one batch
OBS_THERM = 4000    # sweeps from the cold start before sampling, as the stored FP64 reference
                    # (the plaquette plateaus by ~800 sweeps, the U(1) plaquette settles to
                    # 1e-7 only later)
OBS_SWEEPS = 4000   # sampled sweeps
OBS_EVERY = 20      # sweeps between measurements


def observables(chains=OBS_CHAINS, L=OBS_L, therm=OBS_THERM, sweeps=OBS_SWEEPS, every=OBS_EVERY):
    """FP32-mixed against FP64 on the configuration-2 couplings and sweep mix:
    `chains` independent chains per precision (their own Philox keys), cold
    start, `therm` sweeps, then `sweeps` sweeps measured every `every`.  Per
    observable: the mean over chains of each chain's time average, its error
    from the chain-to-chain spread (chains are independent, so
    autocorrelations need no modelling), FP32 - FP64 in standard deviations,
    and the plateau check (first half of the sampling against the second
    half, in standard deviations)."""
    from su2higgs_amd import abi
    from su2higgs_amd.lattice import Batch
    out = {"L": L, "chains": chains, "thermalization_sweeps": therm, "sampled_sweeps": sweeps,
           "every": every}
    # 128 independent FP64 chains of the 48^3 theory at these couplings
    # (written by tests/fp64_reference_48.py)
    with open(os.path.join(ROOT, "tests", "fp64_reference_48.json")) as f:
        ref = json.load(f)
    series = {}
    for prec in (abi.FP32, abi.FP64):
        ps = [workload_params(prec, L=L) for _ in range(chains)]
        for k, p in enumerate(ps):
            p.seed = 7000 + 100 * k + prec
        b = Batch(ps)
        for k in range(chains):
            b.cold_start(k, phi0=1.0)
        b.use_graph(True)
        b.update(therm)
        V = float(abi.volume(ps[0]))
        rows = []
        for _ in range(sweeps // every):
            b.update(every)
            rows.append(b.measure_all()[:chains])
        b.close()
        s_ = np.array(rows)  # [samples, chains, sums]
        series[prec] = {
            "plaquette": 1.0 - s_[:, :, abi.SUM["wilson"]] / (V * ps[0].betasu2) / 3.0,
            "phi2": s_[:, :, abi.SUM["phi2"]] / V,
            "u1_plaquette": 1.0 - s_[:, :, abi.SUM["u1wilson"]] / (V * ps[0].betau1) / 3.0}

    def mean_err(x):
        m = x.mean(0)  # per chain
        return float(m.mean()), float(m.std(ddof=1) / np.sqrt(len(m)))

    for name in ("plaquette", "phi2", "u1_plaquette"):
        e = {}
        for prec, tag in ((abi.FP32, "fp32"), (abi.FP64, "fp64")):
            x = series[prec][name]
            h = len(x) // 2
            m, err = mean_err(x)
            m1, e1 = mean_err(x[:h])
            m2, e2 = mean_err(x[h:])
            e[tag] = [m, err]
            e[tag + "_plateau_sigma"] = (m2 - m1) / float(np.hypot(e1, e2))
        d = e["fp32"][0] - e["fp64"][0]
        se = float(np.hypot(e["fp32"][1], e["fp64"][1]))
        e["fp32_minus_fp64"] = [d, se]
        e["sigma"] = d / se
        if L == 48 and name in ref:
            # each precision against the stored FP64 value of the same theory
            # (512 chains, 4000 sweeps discarded)
            r, re_ = ref[name]
            e["fp64_reference"] = [r, re_]
            for tag in ("fp32", "fp64"):
                m, err = e[tag]
                e[tag + "_vs_reference_sigma"] = (m - r) / float(np.hypot(err, re_))
        out[name] = e
    return out


def run_precision(precision, steps, warmup, timing_steps=100, graph=True):
    """Timed region: `steps` compound sweeps replayed as a hipGraph on the
    library's stream (wall clock between stream synchronisations; with
    graph=False launched kernel by kernel from the host).  Kernel
    region: `timing_steps` sweeps launched one by one with a HIP event pair
    around every launch on the same stream (per-kernel durations)."""
    from su2higgs_amd import abi
    from su2higgs_amd.lattice import Batch
    p = workload_params(precision)
    V = abi.volume(p)
    npass, upd = passes_and_updates(p)
    b = Batch(p)
    b.cold_start(0, phi0=1.0)
    # workgroup chunk per pass kind picked by the autotuner (state restored)
    tune = b.autotune(repeats=50)
    picks = {name: b.chunk(abi.KIND_NAMES.index(name)) for name in tune}
    waves = {name: b.waves(abi.KIND_NAMES.index(name)) for name in tune}
    b.use_graph(graph)
    b.update(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.update(steps)
    b.sync()
    wall = time.perf_counter() - t0
    lp = slot_params(b)
    res = {"precision": precision, "V": V, "passes_per_sweep": npass, "updates_per_sweep_over_V": upd,
           "link_format": {abi.LINKS_FULL: "full", abi.LINKS_PACKED: "packed"}[lp.link_format],
           "hop_cache": b.hop_cache(),
           "autotune_ms_64_128_256": tune, "chunk": picks, "waves": waves,
           "wall_s": wall, "ms_per_sweep": 1e3 * wall / steps,
           "site_updates_per_s": upd * V * steps / wall}
    b.use_graph(False)
    b.set_timing(True)
    ts = max(1, min(steps, timing_steps))
    b.reset_timing()
    e0 = time.perf_counter()
    b.update(ts)
    b.sync()
    res["eager_site_updates_per_s"] = upd * V * ts / (time.perf_counter() - e0)
    ms, n = b.kernel_times()
    b.set_timing(False)
    res["kernel_ms"] = {abi.KIND_NAMES[k]: float(ms[k]) for k in range(abi.NKINDS) if n[k]}
    res["kernel_launches"] = {abi.KIND_NAMES[k]: int(n[k]) for k in range(abi.NKINDS) if n[k]}
    res["timing_sweeps"] = ts
    # average duration of every pass kind of the sweep over 200 back-to-back
    # launches on the library stream: kernel time from the start/stop events
    # each dispatch stamps, and the span per launch including launch gaps
    res["kernel_avg_ms"], res["kernel_span_ms"] = {}, {}
    for name in res["kernel_launches"]:
        k = abi.KIND_NAMES.index(name)
        if k < abi.K_MEASURE:
            res["kernel_avg_ms"][name], res["kernel_span_ms"][name] = b.time_pass(k, 200)
    res["p"] = lp
    b.close()
    return res


def sweep_ms_with_links(precision, links, steps=200, L=96, graph=True):
    """ms per sweep of the workload with the FP32 links in the given
    storage form (autotuned chunks; graph-replayed unless graph=False), and
    its link overrelaxation's average launch and fraction of 8 TB/s on that
    form's algorithmic bytes: the other form beside `value`."""
    from su2higgs_amd import abi
    from su2higgs_amd.lattice import Batch
    p = workload_params(precision, L=L)
    p.link_format = links
    b = Batch(p)
    b.cold_start(0, phi0=1.0)
    b.autotune(repeats=50)
    b.use_graph(graph)
    b.update(20)
    b.sync()
    t0 = time.perf_counter()
    b.update(steps)
    b.sync()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    b.use_graph(False)
    lo_ms, _ = b.time_pass(abi.K_LINK_OVERRELAX, 200)
    byts = algo_words(abi.K_LINK_OVERRELAX, slot_params(b), b.hop_cache()) * abi.volume(p) * 4
    b.close()
    return {"ms_per_step": ms, "link_overrelax_us": lo_ms * 1e3,
            "link_overrelax_algorithmic_MB": byts / 1e6,
            "link_overrelax_frac": byts / (lo_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def roofline(res):
    from su2higgs_amd import abi
    p = res["p"]
    w = 8 if p.precision == abi.FP64 else 4
    kms = res["kernel_ms"]
    kind = max((k for k in kms if k not in ("measure", "aux", "muca")), key=lambda k: kms[k])
    kid = abi.KIND_NAMES.index(kind)
    avg_ms = res["kernel_avg_ms"][kind]
    words = algo_words(kid, p, res["hop_cache"])
    bytes_per_launch = words * res["V"] * w
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    prec = "fp64" if p.precision == abi.FP64 else "fp32"
    packed = res["link_format"] == "packed"
    span_ms = res["kernel_span_ms"][kind]
    return {"bound": "hbm", "kernel": kind, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(kind, "96cubed", prec, packed),
            "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms,
            "avg_span_ms": span_ms,
            "frac_with_launch_gaps": bytes_per_launch / (span_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "words_per_site_per_pass": words}


def kernel_rooflines(res):
    """Every pass kind of the sweep against the HBM peak: algorithmic bytes
    per launch over the kind's average launch duration."""
    from su2higgs_amd import abi
    p = res["p"]
    w = 8 if p.precision == abi.FP64 else 4
    out = {}
    for kind, avg_ms in res["kernel_avg_ms"].items():
        byts = algo_words(abi.KIND_NAMES.index(kind), p, res["hop_cache"]) * res["V"] * w
        ach = byts / (avg_ms * 1e-3) / 1e9
        out[kind] = {"avg_launch_us": avg_ms * 1e3, "algorithmic_MB": byts / 1e6,
                     "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS}
    return out


def roofline_beyond_cache(L=192, precisions=None):
    """Every pass kind of the workload's sweep on an L^3 lattice whose fields
    (19 L^3 words: 537 MB FP32, 1.07 GB FP64 at 192^3) exceed the 256 MB
    Infinity Cache that holds the 96^3 fields: HBM-only rates.  The first
    entry of each precision's table keeps the dominant kernel (link
    overrelaxation) for the summary fields."""
    from su2higgs_amd import abi
    from su2higgs_amd.lattice import Batch
    out = {}
    for prec in precisions or (abi.FP32, abi.FP64):
        p = workload_params(prec, L=L)
        b = Batch(p)
        b.cold_start(0, phi0=1.0)
        b.update(2)
        V = abi.volume(p)
        w = 4 if prec == abi.FP32 else 8
        p = slot_params(b)
        kinds = {}
        for k in (abi.K_LINK_OVERRELAX, abi.K_LINK_HEATBATH, abi.K_U1_METROPOLIS,
                  abi.K_DOUBLET_OVERRELAX, abi.K_DOUBLET_METROPOLIS):
            ms, span = b.time_pass(k, 60)
            byts = algo_words(k, p, b.hop_cache()) * V * w
            ach = byts / (ms * 1e-3) / 1e9
            kinds[abi.KIND_NAMES[k]] = {"avg_launch_us": ms * 1e3, "algorithmic_MB": byts / 1e6,
                                        "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS}
        b.close()
        out["fp32" if prec == abi.FP32 else "fp64"] = kinds
    lo = out["fp32"]["link_overrelax"] if "fp32" in out else None
    res = {"workload": "%d^3 SU(2)xU(1)+Higgs" % L, "kernels": out}
    if lo:
        res.update({"kernel": "link_overrelax", "avg_launch_ms": lo["avg_launch_us"] * 1e-3,
                    "algorithmic_bytes_per_launch": lo["algorithmic_MB"] * 1e6,
                    "achieved": lo["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": lo["frac"]})
    return res


def serving(n, rate=20.0, seed=2026):
    """Scan serving: a seeded Poisson stream (rate req/s) of SU(2)+Higgs scan
    points, L in {24..64}, betaG in [6,12], msq in [-0.3,-0.1], lambda in
    [0.005,0.03], 100..1000 sweeps, every 10th request 64^3 with RK3 flow
    dt 0.01 to t 4.0; FP32-mixed sweeps, admission by HBM footprint,
    continuous batching (su2hip_serve, realtime)."""
    from su2higgs_amd import serve
    reqs = serve.stream(seed, rate, n, flow_every=10)
    cfg = serve.config()
    res, st = serve.serve(cfg, reqs)
    flow_lat = [r.t_done - r.t_arrival for r, q in zip(res, reqs) if q.flow and r.status == 0]
    return {"requests": n, "rate_per_s": rate, "seed": seed, "completed": st["completed"],
            "failed": st["failed"], "p50_s": st["p50"], "p99_s": st["p99"], "mean_s": st["mean"],
            "max_wait_s": st["max_wait"], "makespan_s": st["makespan"],
            "last_arrival_s": reqs[n - 1].arrival if n else 0.0,
            "site_updates_per_s": st["site_updates_per_s"], "busy_frac": st["busy"] / st["makespan"],
            "ticks": st["ticks"], "peak_resident": st["peak_resident"],
            "peak_bytes": st["peak_bytes"], "hbm_cap_bytes": cfg.hbm_budget, "flows": st["flows"],
            "flow_request_p50_s": float(np.median(flow_lat)) if flow_lat else None,
            "tick_sweeps": cfg.tick_sweeps, "precision": "FP32 sweeps, FP64 sums"}


def multicanonical():
    """48^3 SU(2)+adjoint Higgs, multicanonical in Sigmasq with the READONLY
    weight of 200 bins on [0,2] (seeded slopes in [-5,5]), one global check
    per parity half-sweep, 500 + 2000 sweeps, FP32-mixed (FP64 order
    parameter, weight and accept), through the su2hip driver; compared with
    the reference's own chains of the same run (tests/golden)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import muca48  # the comparison with the reference's recorded chains
    if not os.path.exists(muca48.FIXTURE):
        return None
    ours = muca48.run(fp32=True)
    c = muca48.compare(ours, muca48.reference())
    c["wall_s"] = ours["wall_s"]
    c["sweeps_per_s_incl_io"] = 2500 / ours["wall_s"]
    # the same run as 16 chains in one handle, pooled, against the
    # reference's four chains: mean and histogram distance
    op, wall = muca48.run_batch(fp32=True, chains=16)
    c["chains16"] = muca48.compare_chains(op, muca48.reference())
    c["chains16"]["wall_s"] = wall
    return c


def scan_params(precision):
    """The 96 lattices of the scan: volumes drawn from {16^3,24^3,32^3,48^3},
    betaG=8, lambda=0.02, msq spaced across [-0.30,-0.15], per-lattice seeds."""
    from su2higgs_amd.abi import FP32, make_params
    rng = np.random.default_rng(96)
    Ls = rng.choice([16, 24, 32, 48], size=96)
    return [make_params(L=(int(L),) * 3, nhiggs=1, precision=precision, betasu2=8.0,
                        lambda_phi=0.02, msq_phi=-0.30 + 0.15 * k / 95, link_overrelax=4,
                        scalar_sweeps=4, reunitarize=1 if precision == FP32 else 0,
                        seed=1000 + k) for k, L in enumerate(Ls)]


def _scan_rate(ps, steps, graph, tune=True, kinds=()):
    """Site-updates/s of one batch holding every lattice of ps (hot starts,
    autotuned chunks, `steps` sweeps replayed as a hipGraph), and the
    per-kind roofline over the whole batch."""
    from su2higgs_amd import abi
    from su2higgs_amd.lattice import Batch
    b = Batch(ps)
    for i in range(len(ps)):
        b.hot_start(i)
    if tune:
        b.autotune(repeats=20)
    b.use_graph(graph)
    b.update(3)
    b.sync()
    t0 = time.perf_counter()
    b.update(steps)
    b.sync()
    wall = time.perf_counter() - t0
    Vtot = sum(abi.volume(p) for p in ps)
    _, upd = passes_and_updates(ps[0])
    w = 4 if ps[0].precision == abi.FP32 else 8
    rp = slot_params(b)
    kr = {}
    for k in kinds:
        ms, _ = b.time_pass(k, 50)
        byts = algo_words(k, rp, b.hop_cache()) * Vtot * w
        ach = byts / (ms * 1e-3) / 1e9
        kr[abi.KIND_NAMES[k]] = {"avg_launch_us": ms * 1e3, "algorithmic_MB": byts / 1e6,
                                 "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS}
    b.close()
    return upd * Vtot * steps / wall, 1e3 * wall / steps, kr


def batched_scan(steps, graph=True, unbatched_steps=50):
    """The batched scan three ways on the same 96 lattices:
    - FP32-mixed batched: every lattice in one handle, one launch per pass
      for the whole batch (the headline form of the scan);
    - FP64 batched: the same, in the reference's arithmetic (496 MB of
      fields, beyond the 256 MB Infinity Cache);
    - FP32 one lattice at a time: each lattice alone in its own handle, as
      the reference runs one process per lattice (src/main.c:251-379),
      `unbatched_steps` sweeps each, its own hipGraph.
    Site-updates/s = 20 V per sweep summed over the lattices / wall time."""
    from su2higgs_amd import abi
    from su2higgs_amd.abi import FP32, FP64
    kinds = (abi.K_LINK_HEATBATH, abi.K_LINK_OVERRELAX, abi.K_DOUBLET_OVERRELAX,
             abi.K_DOUBLET_METROPOLIS)
    p32, p64 = scan_params(FP32), scan_params(FP64)
    Vtot = sum(abi.volume(p) for p in p32)
    v32, ms32, kr32 = _scan_rate(p32, steps, graph, kinds=kinds)
    log("[bench] batched fp32 %.3e site-updates/s" % v32)
    v64, ms64, kr64 = _scan_rate(p64, max(5, steps // 2), graph, kinds=kinds)
    log("[bench] batched fp64 %.3e site-updates/s" % v64)
    # one lattice at a time: total sweep work over total wall time
    _, upd = passes_and_updates(p32[0])
    tot_upd, tot_wall = 0.0, 0.0
    for p in p32:
        r, ms, _ = _scan_rate([p], unbatched_steps, graph, tune=False)
        tot_upd += upd * abi.volume(p) * unbatched_steps
        tot_wall += ms * 1e-3 * unbatched_steps
    vun = tot_upd / tot_wall
    log("[bench] fp32 one lattice at a time %.3e site-updates/s" % vun)
    return {"lattices": 96, "total_sites": int(Vtot), "sweeps": steps,
            "ms_per_sweep": ms32, "site_updates_per_s": v32, "kernel_roofline": kr32,
            "fp64_value": v64, "fp64_ms_per_sweep": ms64, "fp64_sweeps": max(5, steps // 2),
            "fp64_kernel_roofline": kr64, "fp64_field_MB": Vtot * 16 * 8 / 1e6,
            "unbatched_value": vun, "unbatched_sweeps_per_lattice": unbatched_steps,
            "fp32_batched_over_fp64_batched": v32 / v64,
            "fp32_batched_over_fp32_unbatched": v32 / vun}


REF_CONFIG = """dim 3
L1 {L}
L2 {L}
L3 {L}
reset 1
iterations {it}
n_thermalize 0
interval 1000000
checkpoint 1000000
run_checks 0
resultsfile measure
latticefile lattice
betasu2 9.0
betau1 27.0
r_u1 1
msq -0.20
lambda 0.015
phi0 1.0
algorithm_su2link heatbath
algorithm_u1link metropolis
algorithm_su2doublet overrelax
scalar_sweeps 4
update_doublet 1
update_links 1
multicanonical 0
measure_local 0
"""


def _ref_wall(exe, dirs, it):
    """Wall time of len(dirs) copies of the reference binary run side by side,
    one per directory, each for `it` sweeps of the 96^3 workload."""
    procs = []
    t0 = time.perf_counter()
    for d in dirs:
        with open(os.path.join(d, "config"), "w") as f:
            f.write(REF_CONFIG.format(L=96, it=it))
        for fn in ("lattice", "measure", "labels"):
            try:
                os.remove(os.path.join(d, fn))
            except OSError:
                pass
        procs.append(subprocess.Popen([exe, "config"], cwd=d, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL))
    ok = True
    for pr in procs:
        try:
            ok = pr.wait(timeout=300) == 0 and ok
        except subprocess.TimeoutExpired:
            pr.kill()
            ok = False
    return (time.perf_counter() - t0) if ok else None


def cpu_baseline():
    """The reference's serial CPU build (oracle/_ref/su2_hu, -DNHIGGS=1 -DU1)
    on this host's cores: wall(4 sweeps) - wall(1 sweep) of 96^3, cold start,
    the reference's own mix (heatbath + U(1) Metropolis + 4 OR + 1 Metropolis
    = 11 V variable updates per sweep; it has no link overrelaxation).  One
    copy alone, then one independent copy per core side by side (the MPI
    build cannot run here; independent chains are its best case), whose
    aggregate rate is the value."""
    exe = os.path.join(ROOT, "oracle", "_ref", "su2_hu")
    if not os.path.exists(exe):
        return oracle_baseline()
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    # A one-GPU share of the host is 16 cores, whatever the affinity set
    # shows: the copies run on at most that many.
    ncores = max(1, min(CPU_SHARE, affinity))
    V = 96 ** 3
    dirs = [tempfile.mkdtemp(prefix="su2ref_") for _ in range(ncores)]
    try:
        rates = {}
        for n in sorted({1, ncores}):
            w1, w4 = _ref_wall(exe, dirs[:n], 1), _ref_wall(exe, dirs[:n], 4)
            if w1 is None or w4 is None:
                return oracle_baseline()
            rates[n] = (n * 11 * V / ((w4 - w1) / 3), (w4 - w1) / 3)
    finally:
        for d in dirs:
            shutil.rmtree(d, ignore_errors=True)
    return {"value": rates[ncores][0], "unit": "site-updates/s", "cores": ncores,
            "kind": "reference", "value_one_core": rates[1][0],
            "host_affinity_cores": affinity, "host_cpu_count": os.cpu_count(),
            "cores_cap": "%d: the host share of one GPU" % CPU_SHARE,
            "sample": "reference serial build (gcc -O3 -fcommon -DNHIGGS=1 -DU1), 96^3 cold start, "
                      "wall(4 sweeps)-wall(1 sweep), 11 V variable updates per sweep: %d "
                      "independent copies side by side, %.2f s/sweep each (one copy alone: "
                      "%.2f s/sweep)" % (ncores, rates[ncores][1], rates[1][1])}


def oracle_baseline():
    """Fallback: the FP64 oracle restatement (same mix as the GPU) on 32^3."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    p = workload_params(64, L=32)
    o = orc.Oracle(p)
    o.cold_start(phi0=1.0)
    o.update(1)
    t0 = time.perf_counter()
    o.update(2)
    per = (time.perf_counter() - t0) / 2
    _, upd = passes_and_updates(p)
    return {"value": upd * 32 ** 3 / per, "unit": "site-updates/s", "cores": 1, "kind": "port",
            "sample": "oracle restatement, 32^3, 2 sweeps of the GPU mix (%.2f s/sweep)" % per}


def muca48_weight(bins=200, lo=0.0, hi=2.0, seed=48):
    """The multicanonical workload's READONLY weight: `bins` bins on [lo, hi],
    slopes drawn from U[-5, 5] (seeded), W_0 = 0, W_{i+1} = W_i + s_i * width;
    pos/W with the reference's virtual end bins (src/multicanonical.c:47-66)."""
    rng = np.random.default_rng(seed)
    edges = np.linspace(lo, hi, bins + 1)
    W = np.concatenate([[0.0], np.cumsum(rng.uniform(-5, 5, bins) * (hi - lo) / bins)])
    pos = np.concatenate([[lo - 1000.0 * (hi - lo)], edges])
    Wf = np.concatenate([[W[0]], W])
    return pos, Wf


def profile_set(sweeps=10):
    """Every kernel kind of the sweep at its workload's size, launched kernel
    by kernel (no graphs), for kernel profilers: the 96^3 SU(2)xU(1)+Higgs
    sweep with a measurement after each sweep; the 48^3 SU(2)+adjoint
    multicanonical sweep (Sigmasq, READONLY weight of 200 bins on [0, 2], one
    global check per parity half-sweep: k_muca_accept, k_muca_restore); the
    RK3 flow of a 64^3 SU(2)+Higgs lattice (dt 0.01 to t 0.5, a measurement
    every 10 steps); the 96^3 sweep's pass kinds at 192^3; the 96-lattice
    scan (autotuned and warmed up first, as the batched scan is timed).
    Each in FP32-mixed and FP64.  Returns the wall time of each part."""
    from su2higgs_amd import abi
    from su2higgs_amd.abi import FP32, FP64, make_params
    from su2higgs_amd.lattice import Batch
    out = {}
    for prec in (FP32, FP64):
        tag = "fp32" if prec == FP32 else "fp64"
        t0 = time.perf_counter()
        b = Batch(workload_params(prec))
        b.cold_start(0, phi0=1.0)
        for _ in range(sweeps):
            b.update(1)
            b.measure(0)
        b.sync()
        b.close()
        out["sweep96_" + tag] = time.perf_counter() - t0
        t0 = time.perf_counter()
        p = make_params(L=(48, 48, 48), nhiggs=0, triplet=1, precision=prec, betasu2=10.0,
                        msq_triplet=-0.12, b4=0.01, a2=0.0, link_overrelax=4, scalar_sweeps=4,
                        reunitarize=1 if prec == FP32 else 0, seed=4848)
        b = Batch(p)
        b.cold_start(0, sigma0=0.5)
        m = abi.Muca()
        m.enabled, m.orderparam, m.mode, m.checks_per_sweep, m.bins = 1, abi.SIGMASQ, abi.READONLY, 1, 200
        m.min, m.max, m.wrk_min, m.wrk_max, m.do_acceptance = 0.0, 2.0, 0.0, 2.0, 1
        pos, W = muca48_weight()
        b.update(50)  # into the weight's range
        b.muca_set(0, m, pos, W)
        for _ in range(sweeps):
            b.update(1)
            b.measure(0)
        b.sync()
        b.close()
        out["muca48_" + tag] = time.perf_counter() - t0
        t0 = time.perf_counter()
        p = make_params(L=(64, 64, 64), nhiggs=1, precision=prec, betasu2=9.0, msq_phi=-0.2,
                        lambda_phi=0.02, link_overrelax=4, reunitarize=1 if prec == FP32 else 0,
                        seed=6464)
        b = Batch(p)
        b.cold_start(0, phi0=1.0)
        b.update(20)
        b.flow(0, abi.FLOW_RK3, 0.01, 0.5, 10)
        b.close()
        out["flow64_" + tag] = time.perf_counter() - t0
        t0 = time.perf_counter()
        b = Batch(workload_params(prec, L=192))
        b.cold_start(0, phi0=1.0)
        b.update(2)
        b.sync()
        b.close()
        out["sweep192_" + tag] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ps = scan_params(prec)
        b = Batch(ps)
        for i in range(len(ps)):
            b.hot_start(i)
        # as the batched scan is timed: autotuned chunks and warmed-up
        # fields (every launch here kernel by kernel)
        b.autotune(repeats=20)
        b.update(23)
        b.sync()
        b.close()
        out["scan96_" + tag] = time.perf_counter() - t0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-muca", action="store_true", help="skip the multicanonical run")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the serving stream (0 skips it)")
    ap.add_argument("--no-obs", action="store_true", help="skip the observable comparison")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the timed sweeps kernel by kernel instead of replaying a "
                         "captured hipGraph (for tools that do not follow graph launches)")
    ap.add_argument("--profile-set", action="store_true",
                    help="only launch every kernel kind at its workload's size, kernel by kernel "
                         "(for rocprofv3)")
    ap.add_argument("--probe-bytes", type=int, default=0,
                    help="only run the streaming copy kernel over this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a batch or scan "
              "would shard its lattices across GPUs with no communication; one large "
              "lattice would split into z slabs with a halo plane per pass)",
              file=sys.stderr)
        sys.exit(2)
    from su2higgs_amd import abi
    if args.probe_bytes:
        from su2higgs_amd.lattice import copy_probe
        ms = copy_probe(args.probe_bytes, repeats=max(1, args.steps))
        print(json.dumps({"copy_probe_bytes": args.probe_bytes, "ms": ms,
                          "GBps": 2 * args.probe_bytes / (ms * 1e-3) / 1e9}))
        return
    if args.profile_set:
        print(json.dumps({"profile_set_wall_s": profile_set()}))
        return
    obs = not args.no_obs
    log("[bench] FP32-mixed 96^3, %d warmup + %d timed sweeps" % (args.warmup, args.steps))
    r32 = run_precision(abi.FP32, args.steps, args.warmup, graph=not args.no_graph)
    log("[bench] fp32 %.3f ms/sweep, %.3e site-updates/s" % (r32["ms_per_sweep"], r32["site_updates_per_s"]))
    s64 = max(1, min(args.steps, 200))
    log("[bench] FP64 plain 96^3, %d timed sweeps" % s64)
    r64 = run_precision(abi.FP64, s64, max(1, args.warmup // 2), graph=not args.no_graph)
    log("[bench] fp64 %.3f ms/sweep, %.3e site-updates/s" % (r64["ms_per_sweep"], r64["site_updates_per_s"]))
    from su2higgs_amd.abi import LINKS_FULL, LINKS_PACKED
    other_fmt = LINKS_PACKED if r32["link_format"] == "full" else LINKS_FULL
    other = sweep_ms_with_links(abi.FP32, other_fmt, steps=min(args.steps, 200), graph=not args.no_graph)
    log("[bench] fp32 with %s links: %.3f ms/sweep" % (
        "packed" if other_fmt == LINKS_PACKED else "full", other["ms_per_step"]))
    roof = roofline(r32)
    roof64 = roofline(r64)
    log("[bench] every pass kind beyond the Infinity Cache (192^3)")
    roof_big = roofline_beyond_cache()
    for pr, t in roof_big["kernels"].items():
        log("[bench] 192^3 %s: %s" % (pr, ", ".join("%s %.2f" % (k, v["frac"]) for k, v in t.items())))
    batched = None
    if not args.no_batch:
        log("[bench] batched scan, 96 lattices")
        batched = batched_scan(max(5, 2 * args.steps), graph=not args.no_graph)
        log("[bench] batched %.3e site-updates/s" % batched["site_updates_per_s"])
    served = None
    if args.serve_requests > 0:
        log("[bench] serving %d requests at 20/s" % args.serve_requests)
        served = serving(args.serve_requests)
        log("[bench] serving p50 %.3f s p99 %.3f s, %.3e site-updates/s"
            % (served["p50_s"], served["p99_s"], served["site_updates_per_s"]))
    muca = None
    if not args.no_muca:
        log("[bench] multicanonical 48^3 adjoint, 2500 sweeps")
        muca = multicanonical()
        if muca:
            c16 = muca["chains16"]
            log("[bench] muca 16 chains <R> %.6f +- %.6f vs reference %.6f +- %.6f (z %.2f, "
                "histogram distance %.3f, reference leave-one-out %.3f..%.3f)"
                % (c16["mean"], c16["err"], c16["ref_mean"], c16["ref_err"], c16["z"],
                   c16["hist_tv_distance"], min(c16["ref_leave_one_out_tv"]),
                   max(c16["ref_leave_one_out_tv"])))
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline()
        log("[bench] cpu %.3e site-updates/s (%s)" % (cpu["value"], cpu["kind"]))
    obs_res = None
    if obs:
        log("[bench] observables: %d chains per precision, %d^3" % (OBS_CHAINS, OBS_L))
        obs_res = observables()
        log("[bench] fp32 - fp64: " + ", ".join("%s %.2f sigma" % (k, obs_res[k]["sigma"])
                                                 for k in ("plaquette", "phi2", "u1_plaquette")))
    out = {
        "metric": "site-updates/s/GPU, FP32-mixed batched vs FP64 plain; p50/p99 scan latency",
        "value": r32["site_updates_per_s"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_sweep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: cold start phi=(1,0,0,0), Philox4x32-10 draws",
        "timed_launch": "hipGraph replay" if not args.no_graph else "kernel by kernel",
        "config": {"workload": "96^3 SU(2)xU(1)+fundamental Higgs, betaG=9 betaY=27 msq=-0.20 "
                               "lambda=0.015, sweep = 1 link heatbath + 4 link overrelax + U(1) "
                               "Metropolis + 4 Higgs overrelax + 1 Higgs Metropolis",
                   "L": 96, "passes_per_sweep": r32["passes_per_sweep"],
                   "variable_updates_per_sweep": "%g V" % r32["updates_per_sweep_over_V"],
                   "precision": "FP32 fields/staples, FP64 reductions"},
        "link_format_fp32": r32["link_format"],
        "hop_cache_fp32": r32["hop_cache"], "hop_cache_fp64": r64["hop_cache"],
        "fp32_other_link_format": dict(
            {"link_format": "packed" if other_fmt == LINKS_PACKED else "full",
             "value": r32["updates_per_sweep_over_V"] * r32["V"] / (other["ms_per_step"] * 1e-3)},
            **other),
        "fp64_value": r64["site_updates_per_s"],
        "fp64_ms_per_step": r64["ms_per_sweep"],
        "fp32_over_fp64": r32["site_updates_per_s"] / r64["site_updates_per_s"],
        "eager_value": r32["eager_site_updates_per_s"],
        "fp64_eager_value": r64["eager_site_updates_per_s"],
        "batched_scan": batched,
        "serving": served,
        "multicanonical": muca,
        "observables": obs_res,
        "roofline": roof,
        "roofline_fp64": roof64,
        "roofline_beyond_cache": roof_big,
        "kernel_roofline_fp32": kernel_rooflines(r32),
        "kernel_roofline_fp64": kernel_rooflines(r64),
        "kernel_avg_ms_fp32": r32["kernel_avg_ms"],
        "kernel_avg_ms_fp64": r64["kernel_avg_ms"],
        "chunk_fp32": r32["chunk"],
        "chunk_fp64": r64["chunk"],
        "waves_fp32": r32["waves"],
        "waves_fp64": r64["waves"],
        "autotune_fp32": r32["autotune_ms_64_128_256"],
        "kernel_span_ms_fp32": r32["kernel_span_ms"],
        "kernel_span_ms_fp64": r64["kernel_span_ms"],
        "kernel_ms_fp32": r32["kernel_ms"],
        "kernel_launches_fp32": r32["kernel_launches"],
        "kernel_ms_fp64": r64["kernel_ms"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
