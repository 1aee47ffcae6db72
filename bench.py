#!/usr/bin/env python3
"""Throughput of the flux engine on one MI355X.

Default workload: chisq_4096 — 4096 seeded parameter sets of the white-dwarf
+ Roche-donor binary with per-set star-2 resolution (10k-150k elements),
each scored against its own 2000-time curve with 1% Gaussian noise,
autoscaled. The CPU setup that builds the grids (the reference's CPU
method) and each set's host layout run once before timing and are reported
separately. A step evaluates every set afresh from its raw element grids
resident in HBM: the per-set pack on the GPU (records sorted into patches,
cones, moments, work codes, work items), the flux pass and the FP64
assembly / chi-squared. The metric is chi-squared evaluations per second.
The same line reports the rate with the raw grids copied over PCIe every
step, and the replay rate of a host-packed resident batch.

  python bench.py [--steps K] [--warmup W] [--workload chisq_4096|wd_donor_45k|disc_spot_300k]
  python bench.py --workload surrogate_200k   # learned surrogate: BF16 MFMA training
  python bench.py --workload da_mcmc          # delayed-acceptance vs plain exact MCMC

Prints one JSON line (plus progress lines on stderr).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lcurve-rs_amd"))
HOST_THREADS = 16            # the GPU box's CPU share
os.environ.setdefault("LCURVE_HIP_THREADS", str(HOST_THREADS))

import lcurve_rs as lc  # noqa: E402
from lcurve_rs import workloads as W  # noqa: E402

FLOP_PER_VISIT = 19.0        # FP32 flops per element-visit of the reference formulation
# Work the flux kernel performs, per unit, on the linear limb-darkening star
# path of the chisq / wd_donor workloads: an element-lane pair evaluated one
# by one costs mu (mul + 2 fma = 5), w*m (1), I(m) (fma = 2), accumulate
# (fma = 2) = 10 FP32 flops; a patch-lane summed from its moments costs
# sum w mu (5), sum w mu^2 (11), c1*s1 + c2*s2 (3), accumulate (1) = 20 FP64.
FLOP_PER_PAIR = 10.0
FLOP_PER_MOMENT_LANE = 20.0
VALU_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E peak
PROFILES = os.path.join(ROOT, "profiles")


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def chisq_models(nsets):
    sets = W.chisq_sets(nsets)
    base = W.chisq_base_text()
    texts = [W.model_text({"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4],
                           "nlat2f": int(s[5]), "nlat2c": int(s[5])}, base) for s in sets]
    return texts, sets


def build_chisq_dpack(nsets, chunk, limb, slots=1):
    """The device-packed chisq_4096 batch with its synthetic observations:
    each set's true curve (unscaled) with 1% Gaussian noise."""
    texts, sets = chisq_models(nsets)
    times = W.linspace_times(-0.2, 1.2, 2000)
    log(f"CPU setup and host layout of {nsets} sets (grids, eclipses, brightness; raw grids staged) ...")
    t0 = time.time()
    models = [lc.Model.from_text(x) for x in texts]
    dp = lc.DevicePackBatch(models, times, 0.001, 1, chunk=chunk, limb=limb, slots=slots)
    t_build = time.time() - t0
    st = dp.stats()
    log(f"setup {st['setup_core_s']:.1f} core-s, layout {st['layout_core_s']:.1f} core-s, wall "
        f"{st['build_wall_s']:.1f} s on {st['threads']} threads; batch tables {st['concat_s']:.1f} s, upload "
        f"{st['upload_s']:.2f} s ({st['input_bytes'] / 1e9:.2f} GB); {st['n_valid']}/{st['n_sets']} valid")
    true = dp.run(scale=False)["flux"]
    noise = W.chisq_noise(nsets, len(times))
    obs = true * (1.0 + 0.01 * noise)
    sig = 0.01 * np.abs(true)
    dp.set_obs(obs, sig, np.ones_like(obs))
    info = {"texts": texts, "times": times, "expose": 0.001, "ndiv": 1, "obs": obs, "sig": sig,
            "scale": True, "unit_per_step": nsets, "build_s": t_build,
            "config": {"workload": "chisq_4096", "sets": nsets, "times": len(times),
                       "star2_nlat": "U{89..344}", "star2_elements": "10k-150k", "noise": 0.01,
                       "autoscale": True, "seed": 42}}
    return dp, models, info


def build_workload(name, nsets):
    """Returns (batch, info, cpu_builder) with observations in place."""
    if name == "chisq_4096":
        texts, sets = chisq_models(nsets)
        times = W.linspace_times(-0.2, 1.2, 2000)
        log(f"CPU setup of {nsets} sets (grids, eclipses, brightness) ...")
        t0 = time.time()
        batch = lc.Batch([lc.Model.from_text(x) for x in texts], times, 0.001, 1)
        t_build = time.time() - t0
        st = batch.stats()
        log(f"setup {st['setup_seconds']:.1f} s, upload {st['upload_seconds']:.1f} s, "
            f"{st['n_valid']}/{st['n_sets']} valid, {st['record_bytes'] / 1e9:.2f} GB records")
        true = batch.run(scale=False)["flux"]
        noise = W.chisq_noise(nsets, len(times))
        obs = true * (1.0 + 0.01 * noise)
        sig = 0.01 * np.abs(true)
        batch.set_obs(obs, sig, np.ones_like(obs))
        info = {"texts": texts, "times": times, "expose": 0.001, "ndiv": 1, "obs": obs, "sig": sig,
                "scale": True, "unit_per_step": nsets, "build_s": t_build,
                "config": {"workload": "chisq_4096", "sets": nsets, "times": len(times),
                           "star2_nlat": "U{89..344}", "star2_elements": "10k-150k", "noise": 0.01,
                           "autoscale": True, "seed": 42}}
        return batch, info
    if name == "wd_donor_45k":
        text, times, expose, ndiv = W.wd_donor()
        model = lc.Model.from_text(text)
        batch = lc.Batch(model, times, expose, ndiv)
        info = {"texts": [text], "times": times, "expose": expose, "ndiv": ndiv, "obs": None, "sig": None,
                "scale": False, "unit_per_step": 1, "build_s": 0.0,
                "config": {"workload": "wd_donor_45k", "star1_elements": 5067, "star2_elements": 39846,
                           "times": 500, "ndiv": 1}}
        return batch, info
    if name == "disc_spot_300k":
        text, times, expose, ndiv = W.disc_spot()
        batch = lc.Batch(lc.Model.from_text(text), times, expose, ndiv)
        info = {"texts": [text], "times": times, "expose": expose, "ndiv": ndiv, "obs": None, "sig": None,
                "scale": False, "unit_per_step": 1, "build_s": 0.0,
                "config": {"workload": "disc_spot_300k", "star2_elements": 200517, "disc_elements": 100713,
                           "times": 5000, "ndiv": 5}}
        return batch, info
    raise SystemExit(f"unknown workload {name}")


def parity_sample(batch, info, out, nsample):
    """FP64 oracle on a few sets: max relative flux error and masks."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    texts = info["texts"]
    idx = np.unique(np.linspace(0, len(texts) - 1, min(nsample, len(texts))).astype(int))
    times = info["times"]
    kw = {}
    if info["obs"] is not None:
        kw = dict(flux2d=info["obs"][idx], ferr2d=info["sig"][idx], weight2d=np.ones((len(idx), len(times))))
    ob = pyoracle.OracleBatch([texts[i] for i in idx], times, info["expose"], info["ndiv"],
                              nthreads=HOST_THREADS, **kw)
    chi, flux = ob.flux(scale=info["scale"], nthreads=HOST_THREADS)
    gflux = out["flux"][idx]
    err = float(np.nanmax(np.abs(gflux - flux) / np.abs(flux)))
    cerr = float(np.nanmax(np.abs(out["chisq"][idx] / chi - 1))) if info["scale"] else None
    return err, cerr, len(idx)


def cpu_baseline(info, budget_s):
    """The FP64 oracle (CPU restatement of the reference's Rayon path,
    parallel over sets) on a bounded sample of the same workload: a whole
    number of sets (all times) sized so that one flux pass takes about 2 s,
    one warm-up pass, then the median of five timed passes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    texts = info["texts"]
    times = info["times"]

    def sample(nsamp):
        idx = np.arange(nsamp)
        kw = {}
        if info["obs"] is not None:
            kw = dict(flux2d=info["obs"][idx], ferr2d=info["sig"][idx], weight2d=np.ones((nsamp, len(times))))
        return pyoracle.OracleBatch([texts[i] for i in idx], times, info["expose"], info["ndiv"],
                                    nthreads=HOST_THREADS, **kw)

    def one(ob):
        t0 = time.perf_counter()
        ob.flux(scale=info["scale"], nthreads=HOST_THREADS, want_flux=False)
        return time.perf_counter() - t0

    nsamp = min(len(texts), 4 * HOST_THREADS)
    ob = sample(nsamp)
    dt = one(ob)   # warm-up at this size, also used to size the sample
    while dt < 2.0 and nsamp < len(texts) and 7 * dt * 2 < budget_s:
        nsamp = min(len(texts), nsamp * 2)
        ob = sample(nsamp)
        dt = one(ob)
    runs = [one(ob) for _ in range(5)]
    med = float(np.median(runs))
    per_set_setup = ob.setup_seconds * HOST_THREADS / nsamp
    return {"value": nsamp / med if info["unit_per_step"] > 1 else 1.0 / med,
            "cores": HOST_THREADS, "kind": "port",
            "sample": f"{nsamp} of {len(texts)} sets, all {len(times)} times, flux + chi^2 only "
                      f"(grids already built); FP64 C++ restatement of the reference's CPU path; "
                      f"1 warm-up pass, median of 5",
            "runs_s": [round(r, 4) for r in runs],
            "setup_core_seconds_per_set": round(per_set_setup, 4), "wall_s": round(med, 3)}


def pmc_summary(workload, chunk, limb):
    """HBM traffic and VALU issue figures of the flux pass (work items + limb
    lanes) from the committed rocprofv3 --pmc summary of this workload, when
    it was taken with the same work-item size and limb setting (else None)."""
    path = os.path.join(PROFILES, f"flux_pass_{workload}_pmc.json")
    try:
        with open(path) as f:
            pm = json.load(f)
    except (OSError, ValueError):
        return None
    return pm if pm.get("chunk") == chunk and pm.get("limb_max_part") == limb else None


def roofline(workload, st, ms_flux, limb):
    pairs = st["pairs_evaluated"]
    # moment sums as evaluated, per lane: one per work item for its summed
    # moment patches, one per patch split per lane
    mom_lanes = st["moment_evals"]
    flop = pairs * FLOP_PER_PAIR + mom_lanes * FLOP_PER_MOMENT_LANE
    tflops = flop / (ms_flux * 1e-3) / 1e12
    # Compulsory HBM bytes of one flux launch: element records and patch
    # moments once, lane tables, partial rows written.
    # (+ the limb lanes' partial-row updates, 16 B per lane slot, counted
    # within the partial rows)
    alg_bytes = st["record_bytes"] + st["n_lanes"] * 12 + 8 * 64 * st["n_items"]
    pm = pmc_summary(workload, st["chunk"], limb)
    traffic = pm["hbm_bytes_per_launch"] if pm else None
    ref_flop = st["visits"] * FLOP_PER_VISIT
    return {"bound": "valu", "achieved": round(tflops, 3), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / VALU_PEAK_TFLOPS, 4), "traffic": traffic,
            "traffic_source": (f"profiles/flux_pass_{workload}_pmc.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                               f"at the same work-item size and limb setting)") if pm else None,
            # The reference formulation's work: 19 FP32 flop per element-visit
            # for every (element, sub-phase) pair. The engine evaluates only a
            # fraction of them; the rest are decided exactly by the per-(patch,
            # tile) work codes (skipped as zero or summed from patch moments).
            "reference_formulation": {
                "flop_per_visit": FLOP_PER_VISIT, "visits_per_launch": st["visits"],
                "equivalent_tflop_s": round(ref_flop / (ms_flux * 1e-3) / 1e12, 2),
                "equivalent_frac_of_valu_peak": round(ref_flop / (ms_flux * 1e-3) / 1e12 / VALU_PEAK_TFLOPS, 3),
                "executed_work_frac": round(flop / ref_flop, 5),
                "visits_evaluated_frac": round(pairs / st["visits"], 5),
                "visits_by_moments_frac": round(st["pairs_by_moments"] / st["visits"], 5)},
            # the flux pass: flux_items_kernel, then limb_kernel on the limb lanes
            "kernel": "flux_items_kernel+limb_kernel", "kernel_ms": round(ms_flux, 4),
            "flop_per_evaluated_pair": FLOP_PER_PAIR, "flop_per_moment_patch_lane": FLOP_PER_MOMENT_LANE,
            # the per-pair figure is that of the linear limb-darkening star path; the
            # Claret, disc and spot terms of disc_spot_300k cost more, so there it
            # understates the work done
            "flop_counts_exact": workload != "disc_spot_300k",
            "pairs_evaluated_per_launch": pairs, "moment_evals_per_launch": mom_lanes,
            "pairs_by_moments_per_launch": st["pairs_by_moments"],
            "limb_pairs_per_launch": st["pairs_limb"],
            # (element, lane) slots the limb pass computes, padding included:
            # its lane fill is limb_pairs / limb_slots
            "limb_slots_per_launch": st["limb_slots"],
            "reference_visits_per_launch": st["visits"],
            "valu_issue_frac_pmc": pm.get("valu_issue_frac") if pm else None,
            "hbm_algorithmic_bytes": alg_bytes,
            "hbm_frac": round((traffic or alg_bytes) / (ms_flux * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}


def roofline_dpack(st, ms_flux, ms_pack, pm):
    """Roofline of the flux pass of a device-packed step (its dominant
    part): FP32 work as in roofline(), against the VALU peak; the pack's
    compulsory bytes beside it."""
    pairs, mom_lanes = st["pairs_evaluated"], st["moment_evals"]
    flop = pairs * FLOP_PER_PAIR + mom_lanes * FLOP_PER_MOMENT_LANE
    tflops = flop / (ms_flux * 1e-3) / 1e12
    moms = st["n_patches"] * 20 * 8
    alg_bytes = st["rec_bytes"] + moms + st["n_tiles"] * 64 * 12 + 8 * 64 * st["n_items"]
    traffic = pm["hbm_bytes_per_launch"] if pm else None
    ref_flop = st["visits"] * FLOP_PER_VISIT
    # The pack reads the raw grids and writes the records, cones, moments,
    # codes and items once (the radix sort's passes over the 8-byte
    # key/value pairs on top).
    pack_bytes = st["input_bytes"] + st["rec_bytes"] + moms + st["n_codes"] + 64 * st["n_items"]
    return {"bound": "valu", "achieved": round(tflops, 3), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / VALU_PEAK_TFLOPS, 4), "traffic": traffic,
            "traffic_source": ("profiles/flux_pass_chisq_4096_dpack_pmc.json (rocprofv3 --pmc FETCH_SIZE, "
                               "WRITE_SIZE at the same work-item size)") if pm else None,
            "reference_formulation": {
                "flop_per_visit": FLOP_PER_VISIT, "visits_per_launch": st["visits"],
                "equivalent_tflop_s": round(ref_flop / (ms_flux * 1e-3) / 1e12, 2),
                "equivalent_frac_of_valu_peak": round(ref_flop / (ms_flux * 1e-3) / 1e12 / VALU_PEAK_TFLOPS, 3),
                "executed_work_frac": round(flop / ref_flop, 5),
                "visits_evaluated_frac": round(pairs / st["visits"], 5),
                "visits_by_moments_frac": round(st["pairs_by_moments"] / st["visits"], 5)},
            "kernel": "flux_items_kernel", "kernel_ms": round(ms_flux, 4),
            "flop_per_evaluated_pair": FLOP_PER_PAIR, "flop_per_moment_patch_lane": FLOP_PER_MOMENT_LANE,
            "flop_counts_exact": True,
            "pairs_evaluated_per_launch": pairs, "moment_evals_per_launch": mom_lanes,
            "pairs_by_moments_per_launch": st["pairs_by_moments"],
            "reference_visits_per_launch": st["visits"],
            "valu_issue_frac_pmc": pm.get("valu_issue_frac") if pm else None,
            "hbm_algorithmic_bytes": alg_bytes,
            "hbm_frac": round((traffic or alg_bytes) / (ms_flux * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "pack": {"ms": round(ms_pack, 4), "compulsory_bytes": pack_bytes,
                     "gb_s": round(pack_bytes / (ms_pack * 1e-3) / 1e9, 1),
                     "hbm_frac": round(pack_bytes / (ms_pack * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def bench_chisq(args):
    """chisq_4096: fresh sets from resident raw grids (the value), the
    PCIe-inclusive rate and the host-packed replay beside it."""
    chunk = args.chunk if args.chunk > 0 else 4096
    limb = args.device_limb
    dp, models, info = build_chisq_dpack(args.nsets, chunk, limb, args.dpack_slots)
    scale = info["scale"]
    nsets = info["unit_per_step"]
    out = dp.run(scale=scale)
    st = dp.stats()
    log(f"device pack: {st['n_items']} work items, {st['rec_bytes'] / 1e9:.2f} GB records")
    log("parity sample against the FP64 oracle ...")
    flux_err, chi_err, nchk = parity_sample(None, info, out, args.parity_sets)
    log(f"max rel flux err {flux_err:.3e} over {nchk} sets")
    chk_sets = sorted({0, nsets // 2, nsets - 1})
    mism = 0
    for s_ in chk_sets:
        d = dp.check(s_)
        mism += sum(v for k, v in d.items() if not k.startswith("host_"))
    log(f"device pack vs host pack on sets {chk_sets}: {mism} differing values")

    for _ in range(args.warmup):
        dp.launch(scale=scale)
    lc.sync()  # the warm-up's last flux pass is still running: outside the timed region
    log(f"timing {args.steps} steps (pack + flux + assembly) ...")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dp.launch(scale=scale)
    lc.sync()
    elapsed = time.perf_counter() - t0
    ms_step = 1e3 * elapsed / args.steps
    value = nsets * args.steps / elapsed
    reps = max(3, min(args.steps, 10))
    ms_pack, ms_flux, ms_total = dp.time(scale=scale, reps=reps)
    # The same step with the raw grids and layout tables copied from pinned
    # host memory first (the boundary hands over host buffers).
    npc = max(2, min(args.steps, 5))
    t0 = time.perf_counter()
    for _ in range(npc):
        dp.upload()
        dp.launch(scale=scale)
    lc.sync()
    pcie = nsets * npc / (time.perf_counter() - t0)
    st = dp.stats()
    pm = None
    path = os.path.join(PROFILES, "flux_pass_chisq_4096_dpack_pmc.json")
    try:
        with open(path) as f:
            pm = json.load(f)
        if pm.get("chunk") != chunk or pm.get("limb", 0) != limb:
            pm = None
    except (OSError, ValueError):
        pm = None
    roof = roofline_dpack(st, ms_flux, ms_pack, pm)
    rec = {"metric": "chi2 evals/sec (1 GPU) at <=1e-6 rel flux err vs FP64",
           "value": round(value, 3), "unit": "chi2/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
           "data": "synthetic (seeded parameter sets and noise; no external data)",
           "config": info["config"],
           "step": "every set evaluated afresh from its raw element grids resident in HBM: GPU pack, flux pass, "
                   "FP64 assembly and chi^2" + ("; consecutive steps alternate between two work-table sets on two "
                                               "streams, so one step's pack runs beside the previous step's flux pass"
                                               if args.dpack_slots > 1 else ""),
           "element_visits_per_step": st["visits"],
           "visits_per_s": round(st["visits"] * args.steps / elapsed, 1),
           "reference_equivalent_tflop_s": round(st["visits"] * FLOP_PER_VISIT / (ms_flux * 1e-3) / 1e12, 2),
           "max_rel_flux_err_vs_fp64": flux_err, "parity_sets": nchk, "max_rel_chisq_err_vs_fp64": chi_err,
           "device_pack_vs_host_pack": {"sets": chk_sets, "differing_values": int(mism)},
           "fresh_sets": {"ms_pack": round(ms_pack, 4), "ms_flux": round(ms_flux, 4),
                          "ms_launch_total": round(ms_total, 4),
                          "pcie_inclusive_chi2_per_s": round(pcie, 1),
                          "upload_bytes": st["input_bytes"], "upload_s": round(st["upload_s"], 4)},
           "host_side": {"threads": st["threads"], "setup_core_s": round(st["setup_core_s"], 2),
                         "layout_core_s": round(st["layout_core_s"], 2),
                         "setup_core_s_per_set": round(st["setup_core_s"] / nsets, 5),
                         "layout_core_s_per_set": round(st["layout_core_s"] / nsets, 5),
                         "build_wall_s": round(st["build_wall_s"], 2), "batch_tables_s": round(st["concat_s"], 2)},
           "tuning": {"chunk": chunk, "items": st["n_items"], "limb": limb, "slots": args.dpack_slots,
                      "limb_chunks": st["n_limb_chunks"], "limb_items": st["n_limb_items"],
                      "split_pairs": st["split_pairs"], "split_limb_lanes": st["split_limb_lanes"]},
           "roofline": roof}
    if not args.no_replay:
        log("host-packed resident batch (CPU pack, replayed) ...")
        lc.set_limb(not args.no_limb, args.limb_max_part)
        t0 = time.time()
        batch = lc.Batch(models, info["times"], 0.001, 1)
        rb_build = time.time() - t0
        batch.set_obs(info["obs"], info["sig"], np.ones_like(info["obs"]))
        if not args.no_autotune:
            batch.autotune(scale=scale)
        bst = batch.stats()
        for _ in range(args.warmup):
            batch.launch(scale=scale)
        batch.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.launch(scale=scale)
        batch.sync()
        rel = time.perf_counter() - t0
        rv = nsets * args.steps / rel
        rec["replay_host_packed"] = {
            "value": round(rv, 1), "ms_per_step": round(1e3 * rel / args.steps, 3),
            "step": "flux pass + assembly over tables packed once on the CPU (the same sets replayed)",
            "cpu_setup_and_pack_s": round(bst["setup_seconds"], 2), "build_wall_s": round(rb_build, 2),
            "tuning": {"chunk": bst["chunk"], "items": bst["n_items"],
                       "limb": (0 if args.no_limb else args.limb_max_part) if bst["limb_chunks"] else 0}}
        del batch
    if not args.no_cpu_baseline:
        log("CPU baseline (FP64 oracle, bounded sample) ...")
        rec["cpu_baseline"] = cpu_baseline(info, budget_s=30.0)
        rec["cpu_baseline"]["unit"] = rec["unit"]
        rec["speedup_vs_cpu_flux"] = round(value / rec["cpu_baseline"]["value"], 1)
        rec["speedup_vs_cpu_flux_pcie_inclusive"] = round(pcie / rec["cpu_baseline"]["value"], 1)
        s = rec["cpu_baseline"]["setup_core_seconds_per_set"]
        # End to end, setup included on both sides (HOST_THREADS cores): the
        # GPU path's measured build (parsing, CPU setup, host layout, batch
        # tables, upload) plus one step; the CPU path's setup plus its flux.
        rec["end_to_end"] = {
            "gpu_chi2_per_s_incl_setup_16_cores": round(nsets / (info["build_s"] + nsets / value), 2),
            "cpu_chi2_per_s_incl_setup_16_cores": round(
                1.0 / (s / HOST_THREADS + 1.0 / rec["cpu_baseline"]["value"]), 2)}
    print(json.dumps(rec), flush=True)


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16


def bench_surrogate(args):
    """Surrogate training: curves from the exact GPU method over the
    parameter box, folded to 256 phase bins (log flux / wdwarf); the
    5-4x512-256 GELU MLP trained by the BF16 MFMA kernels; held-out error
    against the exact method; MFMA utilisation of the training GEMMs."""
    from lcurve_rs import mcmc
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    # The disc model's parameter box: q, iangle, r1, t1, t2 over the
    # batched-chi-squared ranges, everything else as in disc_spot_300k, at
    # reduced grids (the CPU setup of every training curve is the cost).
    base = W.disc_box_text()
    n = args.ncurves
    params = mcmc.sample_box(n, 42)
    log(f"generating {n} curves with the exact GPU method ...")
    t0 = time.perf_counter()
    curves = mcmc.make_training_set(base, params, 256, log=log)
    t_gen = time.perf_counter() - t0
    ok = np.all(np.isfinite(curves), axis=1)
    P, C = params[ok], curves[ok]
    nh = max(1024, len(P) // 20)
    # One fit per epoch count on the same curves; the record is the last.
    fits = []
    for ep in [int(v) for v in str(args.epochs).split(",")]:
        s = mcmc.Surrogate(256, seed=1)
        log(f"training on {len(P) - nh} curves, {ep} epochs ...")
        fit = s.fit(P[nh:], C[nh:], epochs=ep, log=log)
        err = s.predict(P[:nh]) - C[:nh]
        ae = np.abs(err)
        fits.append({"epochs": ep, "train_s": round(fit["seconds"], 2), "final_train_loss": float(fit["loss"][-1]),
                     "heldout_rms_log_flux": float(np.sqrt(np.mean(err ** 2))),
                     "heldout_median_abs": float(np.median(ae)), "heldout_p99_abs": float(np.quantile(ae, 0.99)),
                     "heldout_max_abs": float(ae.max())})
    err_poly = s.baseline(P[:nh]) - C[:nh]
    tm = s.net.time_step(4096, reps=50)
    tflops = tm["gemm_flop"] / (tm["ms_gemm"] * 1e-3) / 1e12
    value = 4096.0 / (tm["ms_step"] * 1e-3)
    # The same network's step at larger batches (synthetic rows): how far the
    # GEMMs are from MFMA-bound at batch 4096.
    from lcurve_rs.surrogate import Mlp
    scaling = {}
    for bb in (16384, 65536):
        net = Mlp((5, 512, 512, 512, 512, 256), seed=2, max_batch=bb)
        rng = np.random.default_rng(0)
        net.set_data(rng.uniform(-1, 1, (bb, 5)).astype(np.float32), rng.standard_normal((bb, 256)).astype(np.float32))
        tb = net.time_step(bb, reps=10)
        scaling[str(bb)] = {"ms_step": round(tb["ms_step"], 4), "gemm_tflops": round(tb["gemm_flop"] / (tb["ms_gemm"] * 1e-3) / 1e12, 2)}
        del net
    rec = {"metric": "surrogate training samples/s (1 GPU)", "value": round(value, 1), "unit": "samples/s",
           "n_gpus": 1, "steps": 50, "warmup": 1, "ms_per_step": round(tm["ms_step"], 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "bf16/f32",
           "data": "synthetic: curves from the exact GPU method over the parameter box",
           "config": {"workload": "surrogate_200k", "curves": n, "bins": 256, "mlp": "5-512x4-256 GELU",
                      "batch": 4096, "epochs": fits[-1]["epochs"],
                      "box": {k: list(v) for k, v in zip(mcmc.PARAM_NAMES, mcmc.BOX.tolist())},
                      "model": "disc_spot_300k (disc, rim, bright spot, irradiation, Claret) at reduced grids",
                      "grids": W.DISC_BOX_GRIDS},
           "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": MFMA_BF16_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5), "traffic": None,
                        "kernel": "gemm_kernel", "gemm_ms_per_step": round(tm["ms_gemm"], 4),
                        "gemm_flop_per_step": tm["gemm_flop"], "larger_batches": scaling},
           "heldout_vs_exact": {"curves": int(nh), "rms_log_flux": float(np.sqrt(np.mean(err ** 2))),
                                "median_abs": float(np.median(ae)), "p99_abs": float(np.percentile(ae, 99)),
                                "max_abs": float(ae.max()),
                                "quadratic_baseline_alone_rms": float(np.sqrt(np.mean(err_poly ** 2)))},
           "epoch_sweep": fits,
           "data_generation_s": round(t_gen, 1), "train_s": round(fit["seconds"], 2),
           "final_train_loss": fit["loss"][-1], "failed_setups": int((~ok).sum())}
    if not args.no_cpu_baseline:
        import mlp_oracle
        dims = (5, 512, 512, 512, 512, 256)
        p = s.net.get_params().astype(np.float64)
        xs = s.frame.to_u(P[nh:nh + 256])
        ys = (C[nh:nh + 256] - s.baseline(P[nh:nh + 256])) / s.y_std
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            mlp_oracle.loss_and_grad(p, dims, xs, ys, bf16_mode=False)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        rec["cpu_baseline"] = {"value": round(256 / dt, 1), "unit": "samples/s", "cores": HOST_THREADS,
                               "kind": "port",
                               "sample": f"{reps} x 256-row forward+backward in FP64 numpy (oracle/mlp_oracle.py)"}
    print(json.dumps(rec), flush=True)


DA_NLAT = (40, 20, 40, 40)
DA_EXPOSE, DA_NDIV = 0.001, 5


def bench_da_mcmc(args):
    """Delayed acceptance (surrogate transition) vs plain exact ensemble
    MCMC on a synthetic 3000-point, 0.5%-noise curve of the one-binary
    geometry (exposure 0.001 in phase, smeared over 5 sub-times): a pilot
    run, a surrogate trained on a ball of ``--radius`` posterior standard
    deviations around the pilot posterior (whitened coordinates), predicting
    the curve at the data times, then both samplers from the same
    equilibrated walkers."""
    from lcurve_rs import mcmc
    base = mcmc.surrogate_base_text(DA_NLAT)
    truth = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])
    times = np.linspace(-0.2, 1.2, 3000)
    tg = mcmc.Target(base, truth, times, args.noise, seed=42, expose=DA_EXPOSE, ndiv=DA_NDIV)
    nw = args.walkers
    # A run can be split over several processes: --mcmc-save writes the
    # pilot and the plain chain so far, --mcmc-load resumes from them (the
    # plain chain continues from its last walkers with a fresh seed, which is
    # the same Markov chain).
    if args.mcmc_load:
        st = np.load(args.mcmc_load)
        if st["pilot_tail"].shape[-1] != 5 or st["start"].shape[0] != nw:
            raise SystemExit(f"{args.mcmc_load} holds another walker count")
        tail, start = st["pilot_tail"], st["start"]
        pilot = {"seconds": float(st["pilot_seconds"]), "n_exact": int(st["pilot_n_exact"])}
        pchain = st["plain_chain"]
        pinfo = {k: float(st["plain_" + k]) for k in ("seconds", "n_exact", "accepted", "proposals")}
        log(f"resumed: pilot {tail.shape[0]} states, plain chain {pchain.shape[0]} steps")
    else:
        rng = np.random.default_rng(1)
        p0 = truth * (1.0 + 1e-4 * rng.standard_normal((nw, 5)))
        log(f"pilot: {nw} walkers x {args.pilot_steps} exact steps ...")
        pr = mcmc.ensemble_mcmc(tg, p0, args.pilot_steps, seed=1, log=log)
        tail = pr["chain"][args.pilot_steps // 2:].reshape(-1, 5)
        start = pr["chain"][-1].copy()
        pilot = {"seconds": pr["seconds"], "n_exact": pr["n_exact"]}
        pchain = np.zeros((0, nw, 5))
        pinfo = {"seconds": 0.0, "n_exact": 0.0, "accepted": 0.0, "proposals": 0.0}
    plain_steps = args.steps if args.plain_steps < 0 else args.plain_steps
    if plain_steps:
        x0 = pchain[-1] if len(pchain) else start
        log(f"plain exact MCMC: {plain_steps} steps (after {len(pchain)}) ...")
        r = mcmc.ensemble_mcmc(tg, x0, plain_steps, seed=2 + len(pchain), log=log)
        pchain = np.concatenate([pchain, r["chain"]])
        for k in ("seconds", "n_exact", "accepted", "proposals"):
            pinfo[k] += float(r[k])
    plain = dict(pinfo, chain=pchain)
    if args.mcmc_save:
        np.savez(args.mcmc_save, pilot_tail=tail, start=start, pilot_seconds=pilot["seconds"],
                 pilot_n_exact=pilot["n_exact"], plain_chain=pchain,
                 **{"plain_" + k: v for k, v in pinfo.items()})
        log(f"state saved to {args.mcmc_save}")
    if args.skip_da:
        e, t = mcmc.ess(pchain, burn=0.1)
        print(json.dumps({"da_mcmc_part": "plain", "walkers": nw, "plain_steps": int(len(pchain)),
                          "ess": round(float(e), 1), "tau": t.round(2).tolist(),
                          "seconds": round(pinfo["seconds"], 1), "ess_per_s": round(e / pinfo["seconds"], 3),
                          "pilot_seconds": round(pilot["seconds"], 1)}), flush=True)
        return
    frame = mcmc.Frame.local(tail, args.radius)
    ncurves = args.ncurves
    log(f"training set: {ncurves} curves over the pilot posterior ball of radius {args.radius} sigma ...")
    t0 = time.perf_counter()
    params = frame.sample(ncurves, 7)
    curves = mcmc.make_training_set(base, params, chunk=4000, log=log, times=times, expose=DA_EXPOSE,
                                   ndiv=DA_NDIV)
    t_gen = time.perf_counter() - t0
    ok = np.all(np.isfinite(curves), axis=1)
    P, C = params[ok], curves[ok]
    nh = max(512, len(P) // 20)
    s = mcmc.Surrogate(seed=1, frame=frame, times=times, degree=args.degree, tail=True)
    fit = s.fit(P[nh:], C[nh:], epochs=args.epochs, log=log)
    err = s.predict(P[:nh]) - C[:nh]
    err_poly = s.baseline(P[:nh]) - C[:nh]
    inners = [int(v) for v in str(args.inner).split(",")]
    sweep = []
    for inner in inners:
        log(f"delayed acceptance, {inner} surrogate moves per exact check: {args.steps} steps ...")
        r = mcmc.ensemble_mcmc(tg, start, args.steps, seed=3, surrogate=s, inner_steps=inner, log=log)
        e, t = mcmc.ess(r["chain"], burn=0.1)
        sweep.append({"inner_steps": inner, "ess_per_s": round(e / r["seconds"], 3), "ess": round(float(e), 1),
                      "tau": t.round(2).tolist(), "seconds": round(r["seconds"], 1),
                      "acceptance": round(r["accepted"] / r["proposals"], 3), "n_exact": r["n_exact"]})
        if len(sweep) == 1:
            da = r
    # Log-likelihood error of the surrogate (and of its polynomial alone) over
    # the plain chain's last states: exact values from one more batch.
    probe = plain["chain"][-1]
    lp_e = tg.logp_exact(probe)
    lp_s = tg.logp_surrogate(s, probe)
    lp_poly = -0.5 * mcmc.autoscaled_chisq(np.exp(s.baseline(probe)), tg.y, tg.ferr, tg.w)
    ess_p, tau_p = mcmc.ess(plain["chain"], burn=0.1)
    ess_d, tau_d = mcmc.ess(da["chain"], burn=0.1)
    sp, sd = mcmc.summarize(plain["chain"], 0.1), mcmc.summarize(da["chain"], 0.1)
    sig = 0.5 * (sp[2] - sp[0])
    value = ess_d / da["seconds"]
    base_v = ess_p / plain["seconds"]
    build_s = t_gen + fit["seconds"]
    u_plain = np.sqrt(np.max(np.sum(frame.to_u(plain["chain"][len(plain["chain"]) // 10:].reshape(-1, 5)) ** 2, 1)))
    out_da = float(np.mean(~frame.contains(da["chain"][len(da["chain"]) // 10:].reshape(-1, 5))))
    rec = {"metric": "MCMC effective samples/s (1 GPU), delayed acceptance", "value": round(value, 3),
           "unit": "ESS/s", "n_gpus": 1, "steps": args.steps, "warmup": int(len(tail) * 2 // nw),
           "ms_per_step": round(1e3 * da["seconds"] / args.steps, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32/f64 exact, bf16 surrogate",
           "data": f"synthetic target curve (3000 times, {args.noise:g} relative noise) of the one-binary geometry",
           "config": {"workload": "da_mcmc", "walkers": nw, "steps": args.steps, "inner_steps": inners[0],
                      "free": list(mcmc.PARAM_NAMES), "truth": truth.tolist(), "grid_nlat": list(DA_NLAT),
                      "expose": DA_EXPOSE, "ndiv": DA_NDIV, "surrogate_curves": ncurves, "epochs": args.epochs,
                      "noise": args.noise,
                      "frame": f"ball of {args.radius} pilot sigma, whitened; Gaussian tail outside",
                      "baseline_degree": args.degree, "outputs": "the 3000 data times (smeared)"},
           "plain": {"ess_per_s": round(base_v, 3), "ess": round(float(ess_p), 1), "tau": tau_p.round(2).tolist(),
                     "steps": int(len(plain["chain"])),
                     "seconds": round(plain["seconds"], 1), "n_exact": int(plain["n_exact"]),
                     "acceptance": round(plain["accepted"] / plain["proposals"], 3),
                     "max_radius_in_frame": round(float(u_plain), 3)},
           "delayed_acceptance": {"ess_per_s": round(value, 3), "ess": round(float(ess_d), 1),
                                  "tau": tau_d.round(2).tolist(), "seconds": round(da["seconds"], 1),
                                  "n_exact": da["n_exact"], "acceptance": round(da["accepted"] / da["proposals"], 3),
                                  "moved_to_exact_check": da["stage1_passed"], "fallback_at": da["fallback_at"],
                                  "fraction_outside_ball": round(out_da, 6),
                                  "ess_per_s_incl_surrogate_build": round(ess_d / (da["seconds"] + build_s), 3)},
           "speedup_ess_per_s": round(value / base_v, 2),
           "inner_sweep": [dict(v, speedup=round(v["ess_per_s"] / base_v, 2)) for v in sweep],
           "speedup_incl_surrogate_build": round(ess_d / (da["seconds"] + build_s) / base_v, 2),
           "posterior": {"names": list(mcmc.PARAM_NAMES), "plain_p16_p50_p84": sp.tolist(),
                         "da_p16_p50_p84": sd.tolist(),
                         "max_median_shift_in_sigma": float(np.max(np.abs(sd[1] - sp[1]) / sig)),
                         "width_ratio": (0.5 * (sd[2] - sd[0]) / sig).round(3).tolist()},
           "surrogate": {"heldout_rms_log_flux": float(np.sqrt(np.mean(err ** 2))),
                         "heldout_rms_log_flux_polynomial_only": float(np.sqrt(np.mean(err_poly ** 2))),
                         "loglike_error_spread": float(np.std(lp_s - lp_e)),
                         "loglike_error_spread_polynomial_only": float(np.std(lp_poly - lp_e)),
                         "train_s": round(fit["seconds"], 2), "data_generation_s": round(t_gen, 1),
                         "center": frame.center.tolist()},
           "pilot": {"seconds": round(pilot["seconds"], 1), "n_exact": pilot["n_exact"]}}
    print(json.dumps(rec), flush=True)


def copy_calibration(mib):
    """A streaming kernel of known traffic (torch.neg over float32: reads and
    writes `mib` MiB each), run under `rocprofv3 --pmc FETCH_SIZE` /
    `WRITE_SIZE` to scale those counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"kernel": "torch.neg float32", "read_bytes": n * 4,
                                           "write_bytes": n * 4, "launches": 5}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="chisq_4096",
                    choices=["chisq_4096", "wd_donor_45k", "disc_spot_300k", "surrogate_200k", "da_mcmc"])
    ap.add_argument("--ncurves", type=int, default=None,
                    help="training curves (surrogate_200k: 200000; da_mcmc: 16000)")
    ap.add_argument("--epochs", default=None,
                    help="surrogate training epochs (surrogate_200k: a comma-separated list, one fit each)")
    ap.add_argument("--walkers", type=int, default=64)
    ap.add_argument("--pilot-steps", type=int, default=200)
    ap.add_argument("--inner", default="100,50,25",
                    help="surrogate moves per exact check (da_mcmc); a comma list runs one sampler per value "
                         "after the same plain run, the first giving the reported value")
    ap.add_argument("--radius", type=float, default=6.5,
                    help="da_mcmc surrogate domain: pilot posterior standard deviations (whitened ball)")
    ap.add_argument("--degree", type=int, default=3, help="da_mcmc surrogate polynomial baseline degree")
    ap.add_argument("--noise", type=float, default=0.005, help="relative noise of the da_mcmc target curve")
    ap.add_argument("--plain-steps", type=int, default=-1,
                    help="da_mcmc: plain exact steps in this process (default --steps)")
    ap.add_argument("--skip-da", action="store_true", help="da_mcmc: plain chain only (with --mcmc-save)")
    ap.add_argument("--mcmc-save", default=None, help="da_mcmc: write the pilot and plain chain here (.npz)")
    ap.add_argument("--mcmc-load", default=None, help="da_mcmc: resume from a --mcmc-save file")
    ap.add_argument("--nsets", type=int, default=4096, help="sets of chisq_4096 (4096 is the workload)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--chunk", type=int, default=0,
                    help="fix elements per work item (skips autotuning; used for profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-limb", action="store_true",
                    help="evaluate no-eclipse patches on all lanes of a tile (no limb lanes)")
    ap.add_argument("--dpack-slots", type=int, default=1,
                    help="chisq_4096 fresh sets: work-table sets launches alternate between (2: the pack of one "
                         "launch runs beside the flux pass of the one before)")
    ap.add_argument("--device-limb", type=int, default=0,
                    help="chisq_4096 fresh sets: limb lanes in the device pack (0 off, else the split threshold)")
    ap.add_argument("--limb-max-part", type=int, default=48,
                    help="split a no-eclipse patch per lane when at most this many lanes see it in part")
    ap.add_argument("--parity-sets", type=int, default=8)
    ap.add_argument("--no-replay", action="store_true",
                    help="chisq_4096: skip the host-packed replay batch")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a streaming kernel of known traffic (for PMC scaling)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if not lc.has_hip():
        raise SystemExit("no HIP device: the engine has no CPU path")
    lc.set_device("hip:0")
    if args.workload == "surrogate_200k":
        args.ncurves = args.ncurves or 200000
        args.epochs = args.epochs or "60,600"
        bench_surrogate(args)
        return
    if args.workload == "da_mcmc":
        args.ncurves = args.ncurves or 16000
        args.epochs = int(args.epochs or 200)
        if args.steps == 10:
            args.steps = 400
        bench_da_mcmc(args)
        return
    if args.workload == "chisq_4096":
        bench_chisq(args)
        return
    if args.chunk > 0:
        lc.set_tuning(args.chunk, 0)
    lc.set_limb(not args.no_limb, args.limb_max_part)

    batch, info = build_workload(args.workload, args.nsets)
    scale = info["scale"]
    if not args.no_autotune and args.chunk <= 0:
        log("autotuning chunk size / launch form ...")
        batch.autotune(scale=scale)
    st = batch.stats()
    log(f"tuning: chunk {st['chunk']}, persistent {st['persistent']}, items {st['n_items']}")

    # correctness on a sample against the FP64 oracle
    out = batch.run(scale=scale)
    log("parity sample against the FP64 oracle ...")
    flux_err, chi_err, nchk = parity_sample(batch, info, out, args.parity_sets)
    log(f"max rel flux err {flux_err:.3e} over {nchk} sets")

    for _ in range(args.warmup):
        batch.launch(scale=scale)
    batch.sync()
    log(f"timing {args.steps} steps ...")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch(scale=scale)
    batch.sync()
    elapsed = time.perf_counter() - t0
    ms_step = 1e3 * elapsed / args.steps
    value = info["unit_per_step"] * args.steps / elapsed

    # Dominant kernel, timed with HIP events on its own stream.
    ms_flux, ms_pass = batch.time(scale=scale, reps=max(3, min(args.steps, 10)))
    roof = roofline(args.workload, st, ms_flux,
                    (0 if args.no_limb else args.limb_max_part) if st["limb_chunks"] else 0)

    rec = {"metric": "chi2 evals/sec (1 GPU) at <=1e-6 rel flux err vs FP64" if args.workload == "chisq_4096"
           else "light curves/sec (1 GPU)",
           "value": round(value, 3), "unit": "chi2/s" if args.workload == "chisq_4096" else "curves/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
           "data": "synthetic (seeded parameter sets and noise; no external data)",
           "config": info["config"],
           "element_visits_per_step": st["visits"],
           "visits_per_s": round(st["visits"] * args.steps / elapsed, 1),
           # rate the reference's per-visit arithmetic would need for this throughput
           "reference_equivalent_tflop_s": round(st["visits"] * FLOP_PER_VISIT / (ms_flux * 1e-3) / 1e12, 2),
           "max_rel_flux_err_vs_fp64": flux_err, "parity_sets": nchk,
           "max_rel_chisq_err_vs_fp64": chi_err,
           "cpu_setup_s": round(st["setup_seconds"], 2), "upload_s": round(st["upload_seconds"], 2),
           "tuning": {"chunk": st["chunk"], "persistent": st["persistent"], "items": st["n_items"],
                      "limb": (0 if args.no_limb else args.limb_max_part) if st["limb_chunks"] else 0,
                      "limb_chunks": st["limb_chunks"]},
           "roofline": roof}
    if not args.no_cpu_baseline:
        log("CPU baseline (FP64 oracle, bounded sample) ...")
        rec["cpu_baseline"] = cpu_baseline(info, budget_s=30.0)
        rec["cpu_baseline"]["unit"] = rec["unit"]
        rec["speedup_vs_cpu_flux"] = round(value / rec["cpu_baseline"]["value"], 1)
        # End to end. GPU: this batch's measured build (model parsing, CPU
        # setup and packing on HOST_THREADS threads, upload; wall clock)
        # plus the flux rate. CPU: the oracle's setup (the same setup code)
        # plus its flux rate.
        s = rec["cpu_baseline"]["setup_core_seconds_per_set"]
        if info["unit_per_step"] > 1:
            nsets = info["unit_per_step"]
            build_s = info["build_s"]
            rec["end_to_end"] = {
                "gpu_chi2_per_s_incl_setup_16_cores": round(nsets / (build_s + nsets / value), 2),
                "cpu_chi2_per_s_incl_setup_16_cores": round(
                    1.0 / (s / HOST_THREADS + 1.0 / rec["cpu_baseline"]["value"]), 2)}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
