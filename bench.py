#!/usr/bin/env python3
"""Headline benchmark: link updates/s at 48^4, FP32 bulk vs FP64 plain, on one GPU.

Workload (one "step" = one sweep of the hot path over the whole lattice):
  48^4 compact U(1), Wilson action, beta = 1.0, hot start;
  sweep = 1 Metropolis pass (4 hits, theta' = theta + U[-0.5, 0.5)) + 3
  overrelaxation passes, each pass = 8 checkerboard (parity, mu) launches;
  the plaquette is measured after every sweep (FP64 reduction, on the device),
  by default summed inside the last overrelaxation pass (--plaq fused) from
  the staple pieces and the new angles, or by a separate pass (--plaq separate).
Defaults follow the 48^4 protocol: 500 thermalisation sweeps (--warmup) then
1000 measured, timed sweeps (--steps), for the FP32-bulk path and again for
the FP64 path on the same GPU.  A link update is one link visited by one pass,
so a sweep is 16 V = 84,934,656 link updates.

The JSON line carries:
  value          FP32-bulk link updates/s (whole timed region, inputs resident)
  roofline       dominant FP32 kernel: the launch class of the sweep with the
                 largest time per sweep (Metropolis phase, overrelaxation
                 phase, measuring phase, or a fused (parity, 0) + (parity, 1)
                 pair): algorithmic bytes per launch of that class (phase:
                 4V angles read + V/2 written; pair: 4V read + V written; 4 B
                 each) / mean launch duration from HIP events on the library
                 stream around 32 back-to-back launches of that class in the
                 sweep's shapes (on a scratch copy of the lattice after the
                 timed region), against 8 TB/s; sweep_frac: the whole sweep's
                 algorithmic bytes over the timed region, against 8 TB/s;
                 traffic (HBM bytes per launch) from rocprofv3 FETCH_SIZE /
                 WRITE_SIZE counter files when given with --traffic, else null;
                 stream_reference_GBs: plain streaming kernels moving the same
                 bytes (read only, copy, and the phase's 8:1 read:write mix)
  batched        the 64-replica mix of uneven lattices (8^4 ... 24^3 x 32,
                 beta ~ U[0.9, 1.1], hot/cold starts), one hipGraph per sweep
                 with every replica's plaquette summed inside the sweep: FP32
                 and FP64 aggregate link updates/s, launch classes, roofline
  cpu_baseline   the FP64 host oracle (OpenMP) on a bounded sample of the
                 same sweeps, timed on this host's cores
  fp64 / ratio / plaquette errors as extra fields.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from latticeu1_amd.analysis import gamma_method, jackknife  # noqa: E402

HBM_PEAK_GBS = 8000.0


# launch classes of u1_bench_updates' kernel timing
CLASSES = ("metropolis", "overrelaxation", "overrelaxation_with_plaquette", "metropolis_pair",
           "overrelaxation_pair", "overrelaxation_pair_with_plaquette", "metropolis_pass", "overrelaxation_pass",
           "overrelaxation_pass_with_plaquette")
NCLS = len(CLASSES)
# phases per launch: 1 (phase), 2 (fused pair), 4 (fused parity pass)
PHASES = (1, 1, 1, 2, 2, 2, 4, 4, 4)
KERNEL_NAMES = ("k_phase<float> Metropolis (parity, mu) phase",
                "k_phase<float> overrelaxation (parity, mu) phase",
                "k_phase<float, MEAS> overrelaxation phase summing the plaquette",
                "k_pair<float> fused Metropolis (parity, 0) + (parity, 1)",
                "k_pair<float> fused overrelaxation (parity, 0) + (parity, 1)",
                "k_pair<float, MEAS> fused overrelaxation pair summing the plaquette",
                "k_chain<float> fused Metropolis parity pass (phases 0-3)",
                "k_chain<float> fused overrelaxation parity pass (phases 0-3)",
                "k_chain<float, MEAS> fused overrelaxation parity pass summing the plaquette")
# kernel-trace names of each class (template arguments: phase <Real, MU, KIND,
# BATCHED, SPT, MEAS>, pair <Real, KIND, SPT, MEAS, BATCHED, FORM>, pass <Real,
# KIND, SPT, MEAS>; KIND 0 Metropolis, 1 overrelaxation)
TRACE_PATTERNS = (r"k_phase<float, \d, 0, (false|true), \d, false>",
                  r"k_phase<float, \d, 1, (false|true), \d, false>",
                  r"k_phase<float, \d, 1, (false|true), \d, true>",
                  r"k_pair<float, 0, \d, false(, (false|true)(, \d)?)?>",
                  r"k_pair<float, 1, \d, false(, (false|true)(, \d)?)?>",
                  r"k_pair<float, 1, \d, true(, (false|true)(, \d)?)?>",
                  r"k_chain<float, 0, \d, false>",
                  r"k_chain<float, 1, \d, false>",
                  r"k_chain<float, 1, \d, true>")


def algorithmic_bytes(c, V, s):
    """Compulsory bytes of one launch of class c: every phase reads the whole
    lattice (4V angles) and writes the V/2 links it updates, so a launch of k
    phases moves (4 + k/2) V angles: 4.5 V (phase), 5 V (pair), 6 V (pass)."""
    return (4.0 + 0.5 * PHASES[c]) * V * s


def ensemble_specs(seed=2024, n=64):
    """The batched-ensemble workload: n replicas drawn by seed from 8^4,
    12^3 x 16, 16^4, 16^3 x 32 and 24^3 x 32, beta ~ U[0.90, 1.10], hot or
    cold starts, Philox seeds."""
    rng = np.random.default_rng(seed)
    shapes = [(8, 8, 8, 8), (12, 12, 12, 16), (16, 16, 16, 16), (16, 16, 16, 32), (24, 24, 24, 32)]
    dims = [shapes[i] for i in rng.integers(0, len(shapes), n)]
    betas = rng.uniform(0.90, 1.10, n)
    hot = rng.integers(0, 2, n)
    seeds = rng.integers(1, 2**40, n)
    return dims, betas, hot, seeds


def run_path(u1, prec, dims, beta, seed, warmup, steps, plaq_mode):
    g = u1.geom(dims, beta, 1.0)
    lat = u1.Lattice.cold(g, prec)
    lat.hotstart(seed)
    m = u1.BENCH_METRO
    if warmup > 0:
        lat.bench(1, 3, m, seed, 0, warmup, 0)  # thermalisation (also runs the autotuner)
    ms, kernel_ms, series = lat.bench(1, 3, m, seed, warmup, steps, plaq_mode, kernel_timing=True)
    return lat, ms, kernel_ms, series


def class_table(kms, V, s):
    """Per launch class: mean ms, launches per sweep, algorithmic bytes per
    launch (V: sites the launch covers, over all replicas), achieved GB/s,
    time per sweep and its share of the sum over classes."""
    out = {}
    for c, name in enumerate(CLASSES):
        if kms[NCLS + c] > 0:
            b = algorithmic_bytes(c, V, s)
            t = float(kms[c])
            out[name] = {"mean_launch_ms": round(t, 5), "per_sweep": round(float(kms[NCLS + c]), 3),
                         "bytes_per_launch": b, "achieved_GBs": round(b / (t * 1e-3) / 1e9, 1),
                         "ms_per_sweep": round(t * float(kms[NCLS + c]), 5)}
    tot = sum(v["ms_per_sweep"] for v in out.values())
    for v in out.values():
        v["time_share"] = round(v["ms_per_sweep"] / tot, 4) if tot > 0 else None
    return out


def roofline_of(cls, sweep_ms, real="float"):
    """Roofline object for the launch class with the largest time per sweep,
    on its own algorithmic bytes, plus the sweep-level fraction (real: the
    kernels' element type in the kernel name)."""
    dom = max(cls, key=lambda k: cls[k]["ms_per_sweep"])
    d = cls[dom]
    achieved = d["bytes_per_launch"] / (d["mean_launch_ms"] * 1e-3) / 1e9
    sweep_bytes = sum(v["bytes_per_launch"] * v["per_sweep"] for v in cls.values())
    sweep_gbs = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    return dom, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                 "kernel": KERNEL_NAMES[CLASSES.index(dom)].replace("<float", "<" + real), "class": dom,
                 "bytes_per_launch": d["bytes_per_launch"], "mean_launch_ms": d["mean_launch_ms"],
                 "time_share": d["time_share"], "launch_classes": cls,
                 "sweep_algorithmic_bytes": sweep_bytes, "sweep_algorithmic_GBs": round(sweep_gbs, 1),
                 "sweep_frac": round(sweep_gbs / HBM_PEAK_GBS, 4)}


def counter_traffic(fetch_csv, write_csv, cls_name):
    """HBM bytes per launch of the launch class cls_name from two rocprofv3
    --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what
    the same counter reads for the plain copy kernel k_convert_precision<float,
    float> (Grid_Size elements of 4 B read and 4 B written) in the same run.
    Returns (traffic bytes, fetch factor, write factor)."""
    import csv
    import re

    pat = re.compile(TRACE_PATTERNS[CLASSES.index(cls_name)])

    def scan(path):
        vals, calib = [], []
        with open(path) as f:
            for row in csv.DictReader(f):
                name = row["Kernel_Name"]
                if pat.search(name):
                    vals.append(float(row["Counter_Value"]))
                elif "k_convert_precision<float, float>" in name:
                    calib.append((float(row["Counter_Value"]), 4.0 * float(row["Grid_Size"])))
        if not vals or not calib:
            raise ValueError(f"{path}: no {cls_name} or calibration dispatches")
        kib, nbytes = calib[-1]
        return float(np.mean(vals)), kib * 1024.0 / nbytes

    fetch_kib, ff = scan(fetch_csv)
    write_kib, fw = scan(write_csv)
    return fetch_kib * 1024.0 / ff + write_kib * 1024.0 / fw, ff, fw


def run_batched(u1, steps, warmup, skip_fp64):
    """The 64-replica ensemble in FP32 and FP64: graph sweeps with per-sweep
    plaquettes, timed with HIP events; launch classes over all replicas."""
    dims, betas, hot, seeds = ensemble_specs(2024)
    geoms = [u1.geom(d, b) for d, b in zip(dims, betas)]
    vol = sum(int(np.prod(d)) for d in dims)
    out = {"replicas": len(dims), "sites": vol, "link_updates_per_sweep": 16.0 * vol,
           "shapes": {"x".join(map(str, sh)): int(sum(1 for d in dims if d == sh)) for sh in sorted(set(dims))},
           "beta_range": [round(float(betas.min()), 4), round(float(betas.max()), 4)],
           "hot_starts": int(np.sum(hot))}
    for prec, name in ((u1.F32, "fp32"), (u1.F64, "fp64")):
        if prec == u1.F64 and skip_fp64:
            continue
        print(f"[bench] batched {name}: {len(dims)} replicas, {warmup} thermalisation + {steps} timed graph sweeps",
              file=sys.stderr, flush=True)
        e = u1.Ensemble(geoms, seeds, hot, prec, 1, 3, u1.BENCH_METRO)
        if warmup > 0:
            e.sweep(warmup, graph=True)  # thermalisation (also runs the ensemble tuner)
        e.sweep_plaquette(1)  # builds the measuring graph outside the timed region
        u1.stream_rate(0, 1024, reps=1)  # kernel-trace marker: the timed sweeps follow
        ms, kms, series = e.bench(steps, kernel_timing=True)
        s = 4 if prec == u1.F32 else 8
        cls = class_table(kms, vol, s)
        _, roof = roofline_of(cls, ms / steps, "float" if prec == u1.F32 else "double")
        pm = np.array([[u1.mean_plaquette(r) for r in sw] for sw in series])
        out[name] = {"link_updates_per_s": 16.0 * vol * steps / (ms * 1e-3), "ms_per_sweep": ms / steps,
                     "roofline": roof, "mean_plaquette_over_replicas": float(pm[-1].mean())}
        print(f"[bench] batched {name}: {out[name]['link_updates_per_s']:.4e} link updates/s, {ms / steps:.3f} "
              f"ms/sweep; dominant {roof['class']} frac {roof['frac']}, sweep frac {roof['sweep_frac']}",
              file=sys.stderr, flush=True)
        e.destroy()
    if "fp64" in out:
        out["fp32_over_fp64"] = out["fp32"]["link_updates_per_s"] / out["fp64"]["link_updates_per_s"]
    return out


def cpu_baseline(dims, beta, seed, budget_s):
    """FP64 host oracle (the timed CPU baseline), bounded to about budget_s."""
    from oracle.oracle import Oracle

    orc = Oracle()
    og = orc.geom(dims, beta, 1.0)
    lat = orc.hotstart(og, seed)
    om = orc.metro(1, 4, 0.5)
    V = int(np.prod(dims))
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.update(og, lat, 1, 3, om, seed, n)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 5 or el * (n + 1) / n > 3 * budget_s:
            break
    return {"value": 16.0 * V * n / el, "unit": "link updates/s", "cores": orc.threads(), "kind": "port",
            "sample": f"{n} sweep(s) of the same 48^4 workload (FP64 OpenMP oracle, hot start), {el:.1f} s"}


def run_multilevel(u1, levels, warmup, eager, skip_fp64):
    """The multilevel workload: 24^4, beta = 1, hot start, `warmup` sweeps,
    then MultiLevel(n4 = n2 = levels, k4 = k2 = 1, 1 Metropolis + 3 OR per
    sublattice sweep) per precision, timed on the host around the call (which
    returns after its device work).  A short call first settles the launch
    shapes, and one streaming kernel then marks the start of the timed call
    in a kernel trace."""
    u1.set_graph_capture(not eager)
    out = {"workload": "multilevel 24^4, beta=1.0, n4=n2=%d, k4=k2=1, 1 Metropolis + 3 OR per sublattice sweep, "
                       "analytic multihit, R=1..12" % levels,
           "level2_steps": "graph replays" if not eager else "eager launches"}
    for prec, name in ((u1.F32, "fp32"), (u1.F64, "fp64")):
        if prec == u1.F64 and skip_fp64:
            continue
        g = u1.geom((24, 24, 24, 24), 1.0)
        lat = u1.Lattice.cold(g, prec)
        lat.hotstart(404)
        for s in range(warmup):
            lat.update(1, 3, u1.BENCH_METRO, 404, s)
        lat.multilevel(1, 1, 1, 1, 1, 3, u1.BENCH_METRO, 404, call_id=9)  # settles the launch shapes
        u1.stream_rate(0, 1024, reps=1)  # trace marker
        u1.sync()
        t0 = time.perf_counter()
        c = lat.multilevel(levels, 1, levels, 1, 1, 3, u1.BENCH_METRO, 404, call_id=0)
        el = time.perf_counter() - t0
        out[name] = {"seconds_per_call": el, "C_R": [[float(v.real), float(v.imag)] for v in c]}
        print(f"[bench] multilevel {name}: {el:.3f} s per call; C(1) = {c[0]:.6e}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--dims", type=str, default="48,48,48,48")
    ap.add_argument("--beta", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=20241)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--verbose", action="store_true", help="print every autotuner candidate (stderr)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-batched", action="store_true", help="leave out the 64-replica batched ensemble")
    ap.add_argument("--mode", choices=["headline", "batched", "multilevel"], default="headline",
                    help="headline: the 48^4 line (default); batched: only the 64-replica mix; multilevel: only the "
                         "24^4 multilevel call")
    ap.add_argument("--ml-levels", type=int, default=50, help="multilevel mode: n4 = n2")
    ap.add_argument("--ml-eager", action="store_true",
                    help="multilevel mode: enqueue the level-2 steps eagerly instead of replaying one graph")
    ap.add_argument("--plaq", choices=["fused", "separate"], default="fused",
                    help="plaquette every sweep: summed inside the last overrelaxation pass (fused) or by a "
                         "separate pass over the lattice")
    ap.add_argument("--tunecache", type=str, default="",
                    help="launch-shape cache: loaded if the file exists (no tuning launches), else written after "
                         "tuning")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"), default=None,
                    help="counter_collection.csv files of earlier rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs "
                         "of this script (with --calib-mb): fills roofline.traffic")
    ap.add_argument("--pair-fusion", type=int, choices=[0, 1, 2], default=1,
                    help="fused (parity, 0) + (parity, 1) update launches: 0 never, 1 where the tuner measures "
                         "them faster, 2 always")
    ap.add_argument("--pass-fusion", type=int, choices=[0, 1, 2], default=1,
                    help="fused parity passes (phases 0-3 in one launch): 0 never, 1 where the tuner measures "
                         "them faster, 2 always")
    ap.add_argument("--calib-mb", type=float, default=0.0,
                    help="after the run, copy a buffer of this many MiB once (plain copy kernel of known size, "
                         "used to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding along t is designed but not built)", file=sys.stderr)
        sys.exit(2)
    dims = tuple(int(v) for v in args.dims.split(","))
    V = int(np.prod(dims))
    lu_per_sweep = 16.0 * V

    import latticeu1_amd as u1

    u1.init(0)
    dev = u1.device_name()
    if args.verbose:
        u1.set_verbosity(2)
    u1.set_pair_fusion(args.pair_fusion)
    u1.set_pass_fusion(args.pass_fusion)
    if args.tunecache and os.path.exists(args.tunecache):
        u1.tune_load(args.tunecache)
    if args.mode == "batched":
        print(json.dumps({"mode": "batched", "device": dev,
                          "batched": run_batched(u1, args.steps, args.warmup, args.skip_fp64)}))
        return
    if args.mode == "multilevel":
        print(json.dumps({"mode": "multilevel", "device": dev,
                          "multilevel": run_multilevel(u1, args.ml_levels, args.warmup, args.ml_eager,
                                                       args.skip_fp64)}))
        return

    res = {}
    for prec, name in ((u1.F32, "fp32"), (u1.F64, "fp64")):
        if prec == u1.F64 and args.skip_fp64:
            continue
        print(f"[bench] {name}: {args.warmup} thermalisation + {args.steps} timed sweeps at {dims}", file=sys.stderr,
              flush=True)
        lat, ms, kms, series = run_path(u1, prec, dims, args.beta, args.seed, args.warmup, args.steps,
                                        1 if args.plaq == "fused" else 2)
        pm = np.array([u1.mean_plaquette(r) for r in series])
        mean, err = jackknife(pm)
        _, gerr, tau, _ = gamma_method(pm)
        err = max(err, gerr)  # binned jackknife, or the Gamma-method error if larger
        s = 4 if prec == u1.F32 else 8
        cls = class_table(kms, V, s)
        res[name] = dict(lat=lat, ms=ms, rate=lu_per_sweep * args.steps / (ms * 1e-3), kms=kms, cls=cls, plaq=mean,
                         plaq_err=err, tau=tau, elem=s, series=series)
        print(f"[bench] {name}: {res[name]['rate']:.4e} link updates/s, {ms / args.steps:.3f} ms/sweep, "
              f"<P> = {mean:.8f} +- {err:.1e}; launches (ms x per sweep): "
              + ", ".join(f"{k} {v['mean_launch_ms']:.4f} x {v['per_sweep']:g}" for k, v in cls.items())
              , file=sys.stderr, flush=True)

    f32 = res["fp32"]
    cls = f32["cls"]
    for k, v in cls.items():
        nph = PHASES[CLASSES.index(k)]
        if nph > 1:  # nph phases' worth of per-phase algorithmic bytes
            v["phase_equivalent_GBs"] = round(4.5 * nph * V * f32["elem"] / (v["mean_launch_ms"] * 1e-3) / 1e9, 1)
    dom, roofline = roofline_of(cls, f32["ms"] / args.steps)
    achieved, d = roofline["achieved"], cls[dom]
    roofline["kernel_timing"] = {"back_to_back_launches_per_class": int(f32["kms"][2 * NCLS + 1]),
                                 "ms_per_sweep_sum_of_classes": round(sum(v["ms_per_sweep"] for v in cls.values()), 4),
                                 "ms_per_sweep_timed_region": round(f32["ms"] / args.steps, 4)}

    # streaming references moving the same 36 B x V/2 bytes as one FP32 phase
    # with an ideal pattern: what this HBM delivers in practice
    n = V // 2 - (V // 2) % 8
    ref = {"read": u1.stream_rate(0, n), "copy": u1.stream_rate(1, n), "read8_write1": u1.stream_rate(2, n)}
    roofline["stream_reference_GBs"] = {k: round(v, 1) for k, v in ref.items()}
    roofline["frac_of_read8_write1_stream"] = round(achieved / ref["read8_write1"], 4)

    if args.traffic:
        tb, ff, fw = counter_traffic(args.traffic[0], args.traffic[1], dom)
        roofline["traffic"] = round(tb)
        roofline["traffic_over_algorithmic"] = round(tb / d["bytes_per_launch"], 3)
        roofline["traffic_calibration"] = {"fetch_factor": round(ff, 4), "write_factor": round(fw, 4)}

    extra = {}
    if "fp64" in res:
        f64 = res["fp64"]
        extra["fp64_link_updates_per_s"] = f64["rate"]
        extra["fp32_over_fp64"] = f32["rate"] / f64["rate"]
        extra["fp64_launch_classes"] = f64["cls"]
        d = f32["plaq"] - f64["plaq"]
        comb = math.hypot(f32["plaq_err"], f64["plaq_err"])
        thermalised = args.warmup >= 500
        extra["plaquette"] = {"fp32": [f32["plaq"], f32["plaq_err"]], "fp64": [f64["plaq"], f64["plaq_err"]],
                              "tau_int_sweeps": {"fp32": f32["tau"], "fp64": f64["tau"]},
                              "diff": d, "diff_over_combined_err": d / comb if comb > 0 else None,
                              "thermalised": thermalised,
                              "note": ("one chain per precision after the 500-sweep thermalisation" if thermalised
                                       else f"unthermalised: {args.warmup} thermalisation sweeps (< 500), the "
                                            "means still drift from the hot start; not a statistical comparison")}
        # storage rounding: the final FP64 configuration rounded to binary32
        # and measured by the FP32 plaquette pass
        p64 = f64["lat"].plaquette()
        p32 = f64["lat"].astype(u1.F32).plaquette()
        extra["fp64_config_in_fp32_plaquette_rel_err"] = float(abs(u1.mean_plaquette(p32) - u1.mean_plaquette(p64))
                                                               / abs(u1.mean_plaquette(p64)))
    # the timed quantity: the FP32 sweep's in-sweep plaquette of its final
    # configuration against the FP64 plaquette of the same links widened to
    # binary64 (SS, ST and <P>, largest relative difference)
    pt = np.asarray(f32["series"][-1])
    pw = f32["lat"].astype(u1.F64).plaquette()
    errs = [abs(pt[k] - pw[k]) / abs(pw[k]) for k in (0, 2)]
    errs.append(abs(u1.mean_plaquette(pt) - u1.mean_plaquette(pw)) / abs(u1.mean_plaquette(pw)))
    extra["fixed_config_plaquette_rel_err"] = float(max(errs))
    extra["fixed_config_plaquette"] = {"timed_fp32_in_sweep": [float(v) for v in pt],
                                       "fp64_of_same_links": [float(v) for v in pw],
                                       "rel_err_ss_st_mean": [float(v) for v in errs]}

    if not args.skip_batched:
        extra["batched"] = run_batched(u1, args.steps, args.warmup, args.skip_fp64)

    if args.tunecache and not os.path.exists(args.tunecache):
        u1.tune_save(args.tunecache)

    if args.calib_mb > 0:
        n = int(args.calib_mb * 1048576 / 4)
        src, dst = u1.DeviceArray(n, np.float32), u1.DeviceArray(n, np.float32)
        u1.call("u1_convert_precision", src.ptr, u1.F32, dst.ptr, u1.F32, n)
        u1.sync()
        print(f"[bench] calibration copy: {n * 4} bytes read, {n * 4} bytes written", file=sys.stderr, flush=True)

    cpu = None
    if not args.skip_cpu:
        print(f"[bench] cpu baseline (oracle, ~{args.cpu_budget:.0f} s)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(dims, args.beta, args.seed, args.cpu_budget)

    out = {
        "metric": "link updates/s at 48^4, FP32-bulk vs FP64 same GPU; % HBM roofline; <P> err",
        "value": f32["rate"],
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (hot-start gauge field from Philox seed)",
        "config": {"workload": "U(1) Wilson gauge, 48^4, beta=1.0, hot start; sweep = 1 Metropolis "
                               "(4 hits, delta U[-0.5,0.5)) + 3 overrelaxation; plaquette every sweep",
                   "lattice": "x".join(str(d) for d in dims), "precision": "fp32 angles, fp64 reductions",
                   "plaquette": ("summed inside the last overrelaxation pass" if args.plaq == "fused"
                                 else "separate pass"),
                   "link_updates_per_sweep": lu_per_sweep, "device": dev},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
