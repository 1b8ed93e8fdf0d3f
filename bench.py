#!/usr/bin/env python3
"""Benchmark of the adjoint angle-gather Kirchhoff migration on one MI355X.

A step is one full migration of the workload: trace preparation (mute,
doubint), tuple preparation (table interpolation, gradients, smoothing,
aperture weights) and the migrate kernel over every shot, receiver and image
point, with tables and data already resident in HBM.  The metric counts the
nominal ns*nh*nz*nx image-point x trace updates, as the reference evaluates
every one of them.

Prints one JSON line.  Progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "image-pt×trace updates/s, batched FP16/NN vs plain FP32 GPU at CIG relL2≤1e-3"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"fp32": 16.0, "fp16": 16.0, "bf16": 12.0, "u16": 12.0}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="salt",
                    choices=["salt", "salt_cmp1", "salt_stack", "small", "layers", "surrogate",
                             "datum"])
    ap.add_argument("--grf-models", type=int, default=256,
                    help="surrogate: velocity models of the pretraining set")
    ap.add_argument("--fit-steps", type=int, default=8000,
                    help="surrogate: Adam steps on the workload's own tables")
    ap.add_argument("--pretrain-steps", type=int, default=3000,
                    help="surrogate: Adam steps on the random-model records")
    ap.add_argument("--nn-weights", default=None,
                    help="surrogate: load the network's parameters from this .npy file and "
                         "train nothing")
    ap.add_argument("--save-nn-weights", default=None,
                    help="surrogate: save the trained network's parameters to this .npy file")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16", "u16"])
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--variant", default="auto", choices=["auto", "stream", "resident"])
    ap.add_argument("--cpu-shots", type=int, default=4,
                    help="shots of the workload timed on the CPU oracle (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stream-check", action="store_true",
                    help="skip the stream-kernel run that checks the timed run's CIG")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph-captured replays of the step")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the 16-bit table forms of the default line")
    ap.add_argument("--overlap", type=int, default=0,
                    help="layers: extend layers 2..5 upward by this many depth samples and "
                         "stitch the gathers on the GPU (layercom_cig blend)")
    ap.add_argument("--adj", type=int, default=1, choices=[0, 1],
                    help="0: time adj=0 modelling of the workload's migrated image instead")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain device copy of this many MiB (profiler calibration)")
    ap.add_argument("--gather-probe", type=int, default=0, metavar="LANES_PER_LINE",
                    help="only run the L1 gather-rate probe (kmig_probe_gather) with this "
                         "many lanes per 128-byte line (profiler calibration)")
    ap.add_argument("--record-counters", default=None, metavar="OUT_JSON",
                    help="write the migrate kernel's counter profile from rocprofv3 output "
                         "directories (--trace-dir, --pmc-dirs, --copy-dirs) and exit; no GPU")
    ap.add_argument("--trace-dir", default=None)
    ap.add_argument("--pmc-dirs", nargs="*", default=[])
    ap.add_argument("--copy-dirs", nargs="*", default=[])
    ap.add_argument("--gather-dirs", nargs="*", default=[],
                    help="rocprofv3 counter directories of `bench.py --gather-probe 1`")
    args = ap.parse_args()
    if args.record_counters:
        return record_counters(args)
    if args.gather_probe > 0:
        from kirch_datuming_amd import kmig
        ms, loads = kmig.probe_gather(args.gather_probe, 2000)
        print(json.dumps({"gather_probe_lanes_per_line": args.gather_probe, "ms": ms,
                          "wave_loads": loads}))
        return
    if args.copy_probe > 0:
        # a plain copy kernel of known size: the HBM counters of a profiler run
        # on it give the scale factor for the migrate kernel's counters
        import torch
        n = int(args.copy_probe * 1024 * 1024 // 4)
        x = torch.ones(n, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the shot-sharded "
              "multi-GPU path is a design (partial CIG per GPU + one all-reduce), not built "
              "here.", file=sys.stderr)
        sys.exit(2)
    cores = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))

    import numpy as np

    from kirch_datuming_amd import configs, kmig

    if args.config == "layers":
        return bench_layers(args, cores)
    if args.config == "surrogate":
        return bench_surrogate(args, cores)
    if args.config == "datum":
        return bench_datum(args, cores)
    t0 = time.time()
    log(f"building the {args.config} workload on {cores} host threads")
    if args.config == "salt":
        w = configs.salt(threads=cores)
    elif args.config == "salt_cmp1":
        w = configs.salt(threads=cores, cmp=1)
    elif args.config == "salt_stack":
        w = configs.salt(threads=cores, cmp=1, stack=True)
    else:
        w = configs.small(threads=cores)
    log(f"workload ready in {time.time() - t0:.1f} s: {w.name}, {w.updates:.4e} updates")
    if args.adj == 0:
        return bench_model(args, cores, w)

    kv = dict(w.kv, precision=args.precision, variant=args.variant)
    if args.tile:
        kv["tile"] = args.tile
    ax = w.data_axes
    p = kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    kmig.check_params(p, int(ax[2][0]))
    m = kmig.Migrator(p)
    tu = time.perf_counter()
    m.set_tables(0, w.stab, w.sder, w.s_axes)
    m.set_tables(1, w.rtab, w.rder, w.r_axes)
    m.set_data(w.data, ax)
    upload_s = time.perf_counter() - tu
    log(f"tables and data resident in HBM ({upload_s:.2f} s host-to-device)")
    r = timed_steps(m, args.warmup, args.steps, w.updates)
    st = r["stats"]
    value = r["value"]
    launches = st["migrate_launches"]
    mig_avg = r["migrate_ms"]
    upd_per_launch = w.updates / max(1, launches)
    # the CIG of the timed run itself: finite, and against the stream kernel
    # (the reference's per-point order, pinned on the golden cases) on the
    # same full workload
    cig_timed = m.cig()
    counts = None
    if st["variant"] == 2 and args.precision == "fp32":
        counts = m.count_updates()
        log(f"updates evaluated {counts['evaluated'] / counts['nominal']:.3f}, accepted "
            f"{counts['accepted'] / counts['nominal']:.3f} of nominal")
    graph = None
    if st["variant"] == 2 and not args.no_graph:
        # the same step captured as one hipGraph and replayed
        m.set_graph(True)
        rg = timed_steps(m, max(2, args.warmup), args.steps, w.updates)
        same = bool(np.array_equal(m.cig().view(np.uint32), cig_timed.view(np.uint32)))
        graph = {"value": rg["value"], "ms_per_step": rg["ms_per_step"],
                 "events_ms": rg["events_ms"], "replayed": bool(rg["stats"]["graph"]),
                 "cig_bitwise_equal_to_eager": same, "vs_eager": rg["value"] / value}
        log(f"graph replays: {rg['ms_per_step']:.2f} ms/step ({graph['vs_eager']:.4f} x eager), "
            f"bitwise equal {same}")
    m.close()
    parity = {"timed_run_finite": bool(np.all(np.isfinite(cig_timed))),
              "timed_run_abs_max": float(np.abs(cig_timed).max())}
    if st["variant"] == 2 and not args.no_stream_check:
        log("stream kernel (reference order) on the full workload")
        with kmig.Migrator(kmig.parse_params(dict(kv, variant="stream", precision="fp32"),
                                             int(ax[1][0]), ax[1][1], ax[1][2])) as ms:
            ms.set_tables(0, w.stab, w.sder, w.s_axes)
            ms.set_tables(1, w.rtab, w.rder, w.r_axes)
            ms.set_data(w.data, ax)
            ms.run()
            ms.sync()
            cig_stream = ms.cig()
            t_stream = ms.stats()["ms_total"]
        parity["timed_run_vs_stream_kernel_full"] = rel_l2(cig_timed, cig_stream)
        parity["stream_kernel_ms"] = t_stream
        log(f"timed run vs stream kernel: relL2 {parity['timed_run_vs_stream_kernel_full']:.2e}")
        del cig_stream
    del cig_timed

    cpu = None
    forms = {}
    if not args.no_cpu and args.cpu_shots > 0:
        from oracle import oracle as O
        ns = int(ax[2][0])
        k = args.cpu_shots
        shots = [ns * (j + 1) // (k + 1) for j in range(k)]
        step = shots[1] - shots[0] if k > 1 else 1
        shots = [shots[0] + j * step for j in range(k)]
        sub = w.shot_subset(shots)
        log(f"CPU oracle (float, reference loop order) on shots {shots} with {cores} threads")
        tc = time.perf_counter()
        cig_cpu = O.migrate(w.kv, *sub.args(), precision="f32")[0]
        tcpu = time.perf_counter() - tc
        cpu = {"value": sub.updates / tcpu, "unit": "updates/s", "cores": cores,
               "kind": "port",
               "sample": f"{k} evenly spaced shots ({sub.updates:.3e} updates) of the same "
                         f"workload, {tcpu:.1f} s, OpenMP over image points with the serial "
                         f"per-trace preparation of the reference"}
        log("FP64 oracle on the same sample")
        cig64 = O.migrate(w.kv, *sub.args(), precision="f64")[0]
        # the production path on the sample, two shot groups forced so the
        # partial-CIG sums are in it
        cig_sub = sample_run(kmig, kv, w, sub, split=2)
        parity.update({"sample": f"shots {shots}, 2 shot groups forced",
                       **accuracy(cig_sub, cig64, "gpu"),
                       "rel_l2_cpu_f32_vs_f64": rel_l2(cig_cpu, cig64),
                       "rel_l2_gpu_vs_cpu_f32": rel_l2(cig_sub, cig_cpu)})
        log(f"CPU {cpu['value']:.3e} updates/s; relL2 on the sample: GPU vs f64 "
            f"{parity['rel_l2_gpu_vs_f64']:.2e}, CPU f32 vs f64 "
            f"{parity['rel_l2_cpu_f32_vs_f64']:.2e}")
        # the 16-bit table forms of the same workload: timed the same way,
        # accuracy on the same sample
        if args.precision == "fp32" and not args.no_forms:
            for prec in ("fp16", "bf16"):
                kvp = dict(kv, precision=prec)
                with kmig.Migrator(kmig.parse_params(kvp, int(ax[1][0]), ax[1][1],
                                                     ax[1][2])) as mp:
                    mp.set_tables(0, w.stab, w.sder, w.s_axes)
                    mp.set_tables(1, w.rtab, w.rder, w.r_axes)
                    mp.set_data(w.data, ax)
                    rp = timed_steps(mp, args.warmup, min(args.steps, 10), w.updates)
                cp = sample_run(kmig, kvp, w, sub, split=2)
                forms[prec] = {"value": rp["value"], "ms_per_step": rp["ms_per_step"],
                               "migrate_ms": rp["migrate_ms"],
                               "tuple_prep_ms": rp["tuple_prep_ms"],
                               "trace_prep_ms": rp["trace_prep_ms"],
                               "events_ms": rp["events_ms"], "tile": rp["stats"]["tile"],
                               "vs_fp32": rp["value"] / value, **accuracy(cp, cig64, "gpu")}
                log(f"{prec}: {rp['value']:.3e} updates/s ({forms[prec]['vs_fp32']:.3f} x fp32), "
                    f"relL2 vs f64 {forms[prec]['rel_l2_gpu_vs_f64']:.2e}")
            if args.config.startswith("salt"):
                forms["surrogate"] = surrogate_form(args, w, kv, sub, cig64, value)

    (nt, _, dt), (nh, _, _), (ns, _, _) = ax
    kernel = (f"k_migrate_ph (preset {st['tile']})" if st["variant"] == 2 else "k_migrate")
    roof = roofline(args.config, args.precision, st, mig_avg, upd_per_launch, counts, kernel)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else f"{args.precision} tables, f32 math",
        "data": "synthetic",
        "config": {"workload": w.name, "config": args.config, "nz": int(w.s_axes[0][0]),
                   "nx": int(w.s_axes[1][0]), "ns": int(ns), "nh": int(nh), "nt": int(nt),
                   "ng": int(p.ng), "cmp": int(w.kv.get("cmp", 1)),
                   "updates_per_step": w.updates, "precision": args.precision,
                   "tile": p.tile or f"auto (preset {st['tile']})",
                   "variant": {0: "auto", 1: "stream", 2: "resident"}[st["variant"]]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "breakdown_ms": {"trace_prep": r["trace_prep_ms"], "tuple_prep": r["tuple_prep_ms"],
                         "migrate": mig_avg * launches, "events_total": r["events_ms"]},
        "parity": parity,
        "forms": forms or None,
        "graph": graph,
        "upload_s": upload_s,
        "pcie_inclusive_updates_per_s": w.updates / (r["ms_per_step"] * 1e-3 + upload_s),
    }
    print(json.dumps(out), flush=True)


def surrogate_form(args, w, kv, sub, cig64, value_fp32):
    """The learned-surrogate form of the workload: the network fitted on
    records of the workload's own tables, evaluated in every step's tuple
    preparation on both sides, with the per-slot fallback (a shot or
    receiver position whose worst point is more than 1 ms off keeps its
    tables).  Timed like the FP32 line, accuracy on the same shot sample."""
    import numpy as np

    from kirch_datuming_amd import configs, kmig, nn
    ax = w.data_axes
    (ns, s0, ds), (nr, r0, dr) = w.s_axes[2], w.r_axes[2]
    pos = np.concatenate([s0 + np.arange(ns) * ds, r0 + np.arange(nr) * dr])
    t0 = time.perf_counter()
    rec = configs.table_records(np.concatenate([w.stab, w.rtab]),
                                np.concatenate([w.sder, w.rder]), pos, 4096, 20260707)
    net = nn.Surrogate(**configs.SURROGATE_SALT)
    net.set_data(rec)
    losses = net.train(args.fit_steps)
    log(f"surrogate: {rec.shape[0]} records, {args.fit_steps} steps in "
        f"{time.perf_counter() - t0:.1f} s (loss {losses[-1]:.3g})")

    def migrator(data, axes):
        m = kmig.Migrator(kmig.parse_params(kv, int(axes[1][0]), axes[1][1], axes[1][2]))
        m.set_tables(0, w.stab, w.sder, w.s_axes)
        m.set_tables(1, w.rtab, w.rder, w.r_axes)
        m.set_data(data, axes)
        for side in (0, 1):
            m.set_surrogate(side, net, 1e-3)
        return m

    with migrator(w.data, ax) as m:
        r = timed_steps(m, args.warmup, min(args.steps, 5), w.updates)
        sst = [m.surrogate_stats(k) for k in (0, 1)]
    with migrator(sub.data, sub.data_axes) as m:
        m.run()
        m.sync()
        cig = m.cig()
    pts = sum(q["points"] for q in sst)
    used = sum(q["used"] for q in sst)
    out = {"value": r["value"], "ms_per_step": r["ms_per_step"], "vs_fp32": r["value"] / value_fp32,
           "tuple_prep_ms": r["tuple_prep_ms"], "fallback": "per slot, |dT| > 1 ms",
           "points_from_network_fraction": used / pts if pts else 0.0,
           "max_residual_s": max(q["max_residual"] for q in sst),
           "training_records": int(rec.shape[0]), "training_steps": args.fit_steps,
           **accuracy(cig, cig64, "gpu")}
    net.close()
    log(f"surrogate: {out['value']:.3e} updates/s ({out['vs_fp32']:.3f} x fp32), "
        f"{out['points_from_network_fraction']:.3f} of the points from the network, relL2 vs f64 "
        f"{out['rel_l2_gpu_vs_f64']:.2e}")
    return out


def rel_l2(a, b):
    """sqrt(sum (a-b)^2 / sum b^2), rsf_l2cmp's metric (src/rsf_l2cmp.cpp:9-10)."""
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = float(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2) / den)) if den > 0 else float("nan")


def accuracy(cig, ref, who):
    """relL2 and the worst angle bin (its error norm over the largest bin norm
    of the reference: a relative L2 per bin means nothing for near-empty bins)."""
    import numpy as np
    ref = np.asarray(ref, np.float64)
    bnorm = np.sqrt(np.sum(ref ** 2, axis=(1, 2)))
    berr = np.sqrt(np.sum((np.asarray(cig, np.float64) - ref) ** 2, axis=(1, 2)))
    worst = float(berr.max() / bnorm.max()) if bnorm.max() > 0 else None
    return {f"rel_l2_{who}_vs_f64": rel_l2(cig, ref),
            "worst_angle_bin_err_over_max_bin_norm": worst}


def timed_steps(m, warmup, steps, updates):
    """W untimed and K timed full migrations (run + sync each), wall clock
    around the K, kernel times from the context's HIP events."""
    import numpy as np
    for i in range(warmup):
        m.run()
        m.sync()
        log(f"warmup {i + 1}/{warmup}: {m.stats()['ms_total']:.1f} ms")
    mig, tot, prep, tup = [], [], [], []
    m.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.run()
        m.sync()
        st = m.stats()
        mig.append(st["ms_migrate"] / max(1, st["migrate_launches"]))
        tot.append(st["ms_total"])
        prep.append(st["ms_trace_prep"])
        tup.append(st["ms_tuple_prep"])
    m.sync()
    wall = time.perf_counter() - t0
    log(f"timed {steps} steps in {wall:.3f} s")
    return {"value": updates * steps / wall, "ms_per_step": wall / steps * 1e3,
            "migrate_ms": float(np.mean(mig)), "events_ms": float(np.mean(tot)),
            "trace_prep_ms": float(np.mean(prep)), "tuple_prep_ms": float(np.mean(tup)),
            "stats": st}


def sample_run(kmig, kv, w, sub, split):
    """The production path (autotuned resident preset) on a shot subset, with
    `split` shot groups forced."""
    old = os.environ.get("KMIG_SPLIT")
    os.environ["KMIG_SPLIT"] = str(split)
    try:
        ax = sub.data_axes
        with kmig.Migrator(kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])) as ms:
            ms.set_tables(0, w.stab, w.sder, w.s_axes)
            ms.set_tables(1, w.rtab, w.rder, w.r_axes)
            ms.set_data(sub.data, ax)
            ms.run()
            ms.sync()
            return ms.cig()
    finally:
        if old is None:
            del os.environ["KMIG_SPLIT"]
        else:
            os.environ["KMIG_SPLIT"] = old


# VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane
# instructions per second (a wave64 instruction occupies a SIMD for 4 cycles);
# 2 flops per lane for an FMA give the 78.6 TFLOP/s non-packed FP32 rate.
VALU_PEAK_LANE_INSTR = 256 * 4 * 16 * 2.4e9
# VALU issue cost in units of one 16-lane pass per wave64 instruction (4 cycles
# on a SIMD): FP32 add/sub/mul/fma issue at twice that rate on this GPU
# (157.3 TFLOP/s FP32 vector), the transcendental unit (v_rcp_f32 etc.) at
# half of it; measured 56-60, 35-38 and 19 T lane-instr/s respectively
VALU_COST = {"fast": 0.5, "other": 1.0, "trans": 2.0}
# MI355X: compute units, and XCDs (each counts its own GRBM_GUI_ACTIVE cycles)
N_CU, N_XCD = 256, 8


def migrate_source_sha16():
    """Hash of the migrate kernel's sources (the kernel file and the headers
    compiled into it): a counter profile belongs to one build of them."""
    import hashlib
    h = hashlib.sha256()
    for name in ("kmig_migrate.hip", "kmig_internal.h", "kmig_index.h"):
        with open(os.path.join(ROOT, "kirch_datuming_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def roofline(config, precision, st, mig_ms, updates, counts, kernel):
    """The resident migrate kernel is bound by its tap gathers in the vector
    memory path, not by HBM or VALU issue: its receiver tuples stay in
    registers and culled groups load nothing, so the per-update streamed
    tuple of the bytes model (16 B/update FP32) is not its traffic, and its
    three 8-byte gathers per update keep the texture-data unit busy at the
    rate a pure gather loop reaches (kmig_probe_gather).  `frac` is that busy
    fraction (busy cycles from the rocprofv3 counters of the same kernel
    build under profiles/, launch cycles from the time measured here with
    HIP events) over the probe's; the VALU issue rate of the launch's
    instruction mix and the HBM bytes (FETCH_SIZE and WRITE_SIZE,
    copy-calibrated) are reported beside it, the bytes model's 16 B/update
    as a reuse factor over the HBM bytes.  A profile without the L1 passes
    reports the VALU rate as the bound; one recorded for another build of the
    kernel source is marked stale and not used."""
    bpu = BYTES_PER_UPDATE.get(precision, 16.0)
    model = updates * bpu
    base = {"kernel": kernel, "avg_launch_ms": mig_ms, "updates_per_launch": updates,
            "bytes_model_per_update": bpu}
    if counts:
        base["evaluated_fraction"] = counts["evaluated"] / counts["nominal"]
        base["accepted_fraction"] = counts["accepted"] / counts["nominal"]
    # the profile of the preset that ran, else the workload's default one
    prof = os.path.join(ROOT, "profiles", f"migrate_{config}_{precision}_p{st['tile']}_counters.json")
    if not os.path.exists(prof):
        prof = os.path.join(ROOT, "profiles", f"migrate_{config}_{precision}_counters.json")
    if st["variant"] != 2 or not os.path.exists(prof):
        return {"bound": "hbm", "achieved": model / (mig_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": None, "traffic": None,
                "note": "no counter profile for this kernel and workload", **base}
    with open(prof) as f:
        pc = json.load(f)
    sha = migrate_source_sha16()
    stale = pc.get("source_sha16") != sha or int(pc.get("preset", -1)) != int(st["tile"])
    if stale:
        return {"bound": "valu", "achieved": None, "peak": VALU_PEAK_LANE_INSTR / 1e12,
                "unit": "T lane-instr/s", "frac": None, "traffic": None, "stale": True,
                "profile": os.path.relpath(prof, ROOT),
                "note": "the counter profile was recorded for another kernel build or preset",
                **base}
    valu = pc["sq_insts_valu"] * 64.0
    achieved = valu / (mig_ms * 1e-3)
    # issue passes of the launch's instruction mix: the peak rate of this mix
    # is VALU_PEAK_LANE_INSTR x instructions / passes
    n = pc["sq_insts_valu"]
    fast = sum(pc.get(k, 0.0) for k in ("sq_insts_valu_fma_f32", "sq_insts_valu_add_f32",
                                        "sq_insts_valu_mul_f32"))
    trans = pc.get("sq_insts_valu_trans_f32", 0.0)
    passes = (VALU_COST["fast"] * fast + VALU_COST["trans"] * trans +
              VALU_COST["other"] * (n - fast - trans))
    peak = VALU_PEAK_LANE_INSTR * n / passes
    traffic = pc["hbm_bytes_per_launch"]
    hbm = traffic / (mig_ms * 1e-3) / 1e9
    valu_obj = {"achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "T lane-instr/s",
                "frac": achieved / peak}
    l1 = None
    if "tcp_total_cache_accesses_sum" in pc and "l1_probe_accesses_per_clk_per_cu" in pc:
        # the vector memory pipeline: L1 accesses (a 64-lane 8-byte gather is
        # at least 16, one per 4 lanes) against the probe's peak rate, at the
        # clock the profiled launch ran (its GRBM cycles over its duration)
        cyc = pc["grbm_gui_active"] / N_XCD
        clk = cyc / (pc["avg_launch_ms"] * 1e-3)
        live = mig_ms * 1e-3 * clk  # cycles of the launch timed here
        acc = pc["tcp_total_cache_accesses_sum"]
        busy = pc["td_td_busy_sum"] / N_CU  # texture-data busy cycles per CU
        l1 = {"achieved": busy / live, "peak": pc["l1_probe_td_busy"],
              "unit": "texture-data busy fraction", "frac": busy / live / pc["l1_probe_td_busy"],
              "l1_accesses_per_clk_per_cu": acc / N_CU / live,
              "probe_l1_accesses_per_clk_per_cu": pc["l1_probe_accesses_per_clk_per_cu"],
              "accesses_per_load": acc / pc["sq_insts_vmem_rd"], "clock_ghz": clk / 1e9}
    if l1 is not None:
        return {"bound": "l1", "achieved": l1["achieved"], "peak": l1["peak"],
                "unit": l1["unit"], "frac": l1["frac"],
                "peak_note": "the vector memory path of the tap gathers: the texture-data "
                             "unit's busy fraction over the launch (busy cycles from the "
                             "counter profile, launch cycles from the time measured here) "
                             "against the busy fraction of kmig_probe_gather, a pure loop of "
                             "64-lane 8-byte L1-resident gathers on this GPU; the L1 access "
                             "rate is beside it",
                "l1": l1, "valu": {**valu_obj,
                                   "mix": {"fp32_fast": fast / n, "trans": trans / n,
                                           "other": (n - fast - trans) / n}},
                "traffic": traffic, "stale": False, "profile": os.path.relpath(prof, ROOT),
                "valu_lane_instr_per_launch": valu,
                "valu_instr_per_nominal_update": valu / updates,
                "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": hbm / HBM_PEAK_GBS},
                "reuse_factor_bytes_model_over_traffic": model / traffic,
                "profile_launch_ms": pc.get("avg_launch_ms"), **base}
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
            "unit": "T lane-instr/s", "frac": achieved / peak,
            "peak_note": "VALU issue peak of this launch's instruction mix: FP32 "
                         "add/sub/mul/fma at 2x, transcendentals at 0.5x the 39.3 T "
                         "lane-instr/s of the other VALU instructions",
            "mix": {"fp32_fast": fast / n, "trans": trans / n,
                    "other": (n - fast - trans) / n},
            "frac_unweighted": achieved / VALU_PEAK_LANE_INSTR,
            "traffic": traffic, "stale": False, "profile": os.path.relpath(prof, ROOT),
            "valu_lane_instr_per_launch": valu,
            "valu_instr_per_nominal_update": valu / updates,
            "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm / HBM_PEAK_GBS},
            "reuse_factor_bytes_model_over_traffic": model / traffic,
            "profile_launch_ms": pc.get("avg_launch_ms"), **base}


def record_counters(args):
    """Counter profile of the production migrate kernel (k_migrate_ph, not the
    autotuner's Probe instances) from rocprofv3 CSV output: the kernel-trace
    statistics (--trace-dir), the counter passes (--pmc-dirs; per-dispatch
    values averaged over the kernel's dispatches) and a plain device copy of
    known size read and written (--copy-dirs, one FETCH_SIZE and one
    WRITE_SIZE pass of `bench.py --copy-probe MiB`), which scales the HBM
    counters to bytes."""
    import csv
    import glob

    def rows(d, suffix):
        out = []
        for f in sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)):
            with open(f) as fh:
                out += list(csv.DictReader(fh))
        return out

    def production(name):
        # k_migrate_ph<TZ, TX, R, G, Probe, Fmt, Cols, Bin>: Probe is the fifth
        if "k_migrate_ph<" not in name:
            return False
        targs = [t.strip() for t in name.split("<", 1)[1].split(">", 1)[0].split(",")]
        return len(targs) >= 5 and targs[4] == "false"

    def counters(d, want):
        per = {}
        kname = None
        for r in rows(d, "counter_collection.csv"):
            if not want(r["Kernel_Name"]):
                continue
            kname = r["Kernel_Name"]
            per.setdefault(r["Dispatch_Id"], {}).setdefault(r["Counter_Name"], 0.0)
            per[r["Dispatch_Id"]][r["Counter_Name"]] += float(r["Counter_Value"])
        if not per:
            raise SystemExit(f"no dispatch of the kernel in {d}")
        names = set(c for v in per.values() for c in v)
        return kname, {c: sum(v.get(c, 0.0) for v in per.values()) / len(per) for c in names}

    prof = {}
    kname = None
    for d in args.pmc_dirs:
        k, c = counters(d, production)
        kname = kname or k
        for name, v in c.items():
            prof[name.lower()] = v
    calls, avg_ns = 0, 0.0
    if args.trace_dir:
        for r in rows(args.trace_dir, "kernel_stats.csv"):
            if production(r["Name"]):
                calls = int(r["Calls"])
                avg_ns = float(r["AverageNs"])
    factors = {}
    for d in args.copy_dirs:
        _, c = counters(d, lambda n: "copyBuffer" in n)
        nbytes = args.copy_probe * 1024 * 1024 if args.copy_probe else 1024.0 ** 3
        for key in ("FETCH_SIZE", "WRITE_SIZE"):
            if key in c:
                factors[key] = nbytes / (c[key] * 1024.0)
    targs = [t.strip() for t in kname.split("<", 1)[1].split(">", 1)[0].split(",")]
    out = {"kernel": "k_migrate_ph<" + ", ".join(targs) + ">", "workload": args.config,
           "precision": args.precision, "preset": args.tile,
           "source_sha16": migrate_source_sha16(), "avg_launch_ms": avg_ns * 1e-6,
           "trace_calls": calls, **{k: v for k, v in sorted(prof.items())
                                    if k.startswith(("sq_", "tcp_", "td_", "ta_", "grbm_"))}}
    # the L1 gather probe's rate on this GPU: accesses and texture-data busy
    # cycles per CU per cycle (GRBM_GUI_ACTIVE counts the cycles of each of
    # the XCDs)
    for d in args.gather_dirs:
        _, c = counters(d, lambda n: "k_probe_gather" in n)
        cyc = c["GRBM_GUI_ACTIVE"] / N_XCD
        out["l1_probe_accesses_per_clk_per_cu"] = c["TCP_TOTAL_CACHE_ACCESSES_sum"] / N_CU / cyc
        out["l1_probe_td_busy"] = c["TD_TD_BUSY_sum"] / N_CU / cyc
    if "fetch_size" in prof and "write_size" in prof:
        ff = round(factors.get("FETCH_SIZE", 1.0), 2)
        wf = round(factors.get("WRITE_SIZE", 1.0), 2)
        out.update({"fetch_kib_raw": prof["fetch_size"], "fetch_factor": ff,
                    "write_kib_raw": prof["write_size"], "write_factor": wf,
                    "hbm_bytes_per_launch": 1024.0 * (prof["fetch_size"] * ff +
                                                      prof["write_size"] * wf)})
    with open(args.record_counters, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def bench_model(args, cores, w):
    """adj=0: one step models the workload's data from its migrated image (the
    CIG of one migration of the workload's data); the steps run kmig_model,
    which takes the image from host memory and returns the data there, so the
    kernel time (spray + double integration, HIP events) gives the value and
    the wall time the PCIe-inclusive rate."""
    import numpy as np

    from kirch_datuming_amd import kmig

    kv = {k: v for k, v in w.kv.items() if not k.startswith(("direct_mute", "diff"))}
    kv = dict(kv, variant=args.variant)
    if args.tile:
        kv["tile"] = args.tile
    ax = w.data_axes
    img = kmig.migrate(kv, w.data, ax, w.stab, w.sder, w.s_axes, w.rtab, w.rder, w.r_axes)[0]
    log(f"image: the workload's CIG {img.shape}")
    p = kmig.parse_params(dict(kv, adj=0), int(ax[1][0]), ax[1][1], ax[1][2])
    m = kmig.Migrator(p)
    m.set_tables(0, w.stab, w.sder, w.s_axes)
    m.set_tables(1, w.rtab, w.rder, w.r_axes)
    for i in range(args.warmup):
        m.model(img, ax)
        log(f"warmup {i + 1}/{args.warmup}: {m.stats()['ms_total']:.1f} ms")
    ker_ms = []
    tstart = time.perf_counter()
    for _ in range(args.steps):
        data = m.model(img, ax)
        ker_ms.append(m.stats()["ms_total"])
    wall = time.perf_counter() - tstart
    st = m.stats()
    m.close()
    ker = float(np.mean(ker_ms))
    value = w.updates / (ker * 1e-3)
    log(f"timed {args.steps} steps: kernels {ker:.1f} ms, wall {wall / args.steps * 1e3:.1f} ms")

    cpu = parity = None
    if not args.no_cpu and args.cpu_shots > 0:
        from oracle import oracle as O
        ns = int(ax[2][0])
        k = 1  # the oracle models on one thread: one shot is 1e8 .. 2e8 updates
        shots = [ns // 2]
        sub = w.shot_subset(shots)
        sa = sub.args()
        log(f"CPU oracle modelling (float) of shots {shots}")
        tc = time.perf_counter()
        d_cpu = O.model(kv, img, sa[1], *sa[2:], precision="f32")
        tcpu = time.perf_counter() - tc
        cpu = {"value": sub.updates / tcpu, "unit": "updates/s", "cores": 1, "kind": "port",
               "sample": f"{k} shot ({sub.updates:.3e} updates), {tcpu:.1f} s, "
                         "the reference's single-thread point order"}
        d_gpu = kmig.model(kv, img, sa[1], *sa[2:])
        log("FP64 oracle modelling of the same shots")
        d64 = O.model(kv, img, sa[1], *sa[2:], precision="f64")
        parity = {"sample": f"shots {shots}", "rel_l2_gpu_vs_f64": O.rel_l2(d_gpu, d64),
                  "rel_l2_cpu_f32_vs_f64": O.rel_l2(d_cpu, d64)}
        log(f"relL2 GPU vs f64 {parity['rel_l2_gpu_vs_f64']:.2e}, CPU f32 vs f64 "
            f"{parity['rel_l2_cpu_f32_vs_f64']:.2e}")
    (nt, _, _), (nh, _, _), (ns, _, _) = ax
    kern = "k_model_ph" if st["variant"] == 2 else "k_migrate<Model>"
    out = {
        "metric": "adj=0 modelling updates/s (image point x trace)",
        "value": value, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ker, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name + ", adj=0 from its migrated CIG", "config": args.config,
                   "ns": int(ns), "nh": int(nh), "nt": int(nt), "ng": int(p.ng),
                   "updates_per_step": w.updates, "kernel": f"{kern} (tile {st['tile']})",
                   "variant": {0: "auto", 1: "stream", 2: "resident"}[st["variant"]]},
        "roofline": None,
        "cpu_baseline": cpu,
        "parity": parity,
        "pcie_inclusive_updates_per_s": w.updates * args.steps / wall,
        "data_abs_max": float(np.abs(data).max()),
    }
    print(json.dumps(out), flush=True)


FP32_PEAK_TFLOPS = 157.3


def bench_datum(args, cores):
    """Frequency-domain datuming of the workflow's datuming step (one step =
    forward transforms, both operators with their tables, inverse transforms,
    data resident in HBM).  The operators are complex multiply-adds (8 flops
    each) on the vector FP32 units."""
    import numpy as np

    from kirch_datuming_amd import configs, dat

    t0 = time.time()
    log(f"building the datuming workload on {cores} host threads")
    w = configs.datum(threads=cores)
    log(f"workload ready in {time.time() - t0:.1f} s: {w.name}")
    cfg = dat.config(w.kv, w.datum)
    dat.check(cfg)
    dm = dat.Datum(cfg)
    dm.set_green(0, w.green, *w.pos_axis)
    dm.set_green(1, w.green, *w.pos_axis)
    tu = time.perf_counter()
    dm.set_data(w.data, w.data_axes)
    upload_s = time.perf_counter() - tu
    for i in range(args.warmup):
        dm.run()
        dm.sync()
        log(f"warmup {i + 1}/{args.warmup}: {dm.stats()['ms_total']:.1f} ms")
    rec, fwd, inv, tab, rcv, src, bfr, bfs = [], [], [], [], [], [], [], []
    tstart = time.perf_counter()
    for _ in range(args.steps):
        dm.run()
        dm.sync()
        st = dm.stats()
        rec.append(st["ms_total"])
        fwd.append(st["ms_forward"])
        inv.append(st["ms_inverse"])
        tab.append(st["ms_tables"])
        rcv.append(st["ms_receiver"])
        src.append(st["ms_source"])
        bfr.append(st["ms_bf_receiver"])
        bfs.append(st["ms_bf_source"])
    wall = time.perf_counter() - tstart
    out = dm.output()
    dm.close()
    ms_step = wall / args.steps * 1e3
    value = st["cmacs"] * args.steps / wall
    log(f"timed {args.steps} steps: {ms_step:.1f} ms/step, nfft {st['nfft']}, "
        f"{st['cmacs']:.3e} complex MACs per step; butterfly "
        f"{st['bf_panels_receiver']}+{st['bf_panels_source']} panels, "
        f"{np.mean(bfr):.1f}+{np.mean(bfs):.1f} ms")
    rcv_ms = float(np.mean(rcv))
    band_ms = rcv_ms - float(np.mean(bfr))  # the banded kernel alone
    achieved = 8.0 * st["pairs_receiver"] * st["nw"] / (band_ms * 1e-3) / 1e12

    cpu = parity = None
    if not args.no_cpu:
        from oracle import datum as D
        from oracle import oracle as O
        sub = w.corner(96)
        bfk = "butterfly" if sub.kv.get("use_bf", 1) else "direct path"
        log(f"CPU oracle (float, {bfk}) on a 96 x 96 corner of the cube")
        tc = time.perf_counter()
        ref32 = D.datum(sub.kv, sub.data, sub.data_axes, sub.green, sub.pos_axis, sub.green,
                        sub.pos_axis, sub.datum)
        tcpu = time.perf_counter() - tc
        with dat.Datum(cfg) as ds:
            ds.set_green(0, sub.green, *sub.pos_axis)
            ds.set_green(1, sub.green, *sub.pos_axis)
            ds.set_data(sub.data, sub.data_axes)
            ds.run()
            ds.sync()
            gpu_sub = ds.output()
            sst = ds.stats()
        gpu_exact = dat.datum(dict(sub.kv, use_bf=0), sub.data, sub.data_axes, sub.green,
                              sub.pos_axis, sub.green, sub.pos_axis, sub.datum)
        log("FP64 oracle (direct path) on the same corner")
        ref64 = D.datum(dict(sub.kv, use_bf=0), sub.data, sub.data_axes, sub.green,
                        sub.pos_axis, sub.green, sub.pos_axis, sub.datum, precision="f64")
        cpu = {"value": sst["cmacs"] / tcpu, "unit": "complex MAC/s", "cores": cores,
               "kind": "port",
               "sample": f"96 shots x 96 receivers of the same cube, tables and keys "
                         f"({sst['cmacs']:.3e} complex MACs, nfft {sst['nfft']}), {tcpu:.1f} s, "
                         f"the reference's {bfk} (use_bf={int(sub.kv.get('use_bf', 1))}) "
                         f"with OpenMP over frequencies"}
        parity = {"sample": "96 x 96 corner",
                  "rel_l2_gpu_vs_oracle": O.rel_l2(gpu_sub, ref32),
                  "rel_l2_gpu_exact_vs_f64_direct": O.rel_l2(gpu_exact, ref64),
                  "butterfly_effect": O.rel_l2(gpu_sub, gpu_exact),
                  "butterfly_panels": [sst["bf_panels_receiver"], sst["bf_panels_source"]]}
        log(f"CPU {cpu['value']:.3e} complex MAC/s; relL2 GPU vs oracle "
            f"{parity['rel_l2_gpu_vs_oracle']:.2e}, exact GPU vs f64 direct "
            f"{parity['rel_l2_gpu_exact_vs_f64_direct']:.2e}, butterfly effect "
            f"{parity['butterfly_effect']:.2e}")
    (nt, _, _), (nh, _, _), (ns, _, _) = w.data_axes
    line = {
        "metric": "datuming operator complex MACs/s (kirchdat2d_freq)",
        "value": value, "unit": "complex MAC/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (complex)", "data": "synthetic",
        "config": {"workload": w.name, "config": "datum", "ns": int(ns), "nh": int(nh),
                   "nt": int(nt), "nfft": st["nfft"], "keys": w.kv,
                   "pairs_receiver": st["pairs_receiver"], "pairs_source": st["pairs_source"],
                   "kernel": "k_band<8,4>" if st["banded"] else "k_pairs",
                   "butterfly": {"on": st["butterfly"],
                                 "panels": [st["bf_panels_receiver"], st["bf_panels_source"]],
                                 "items": [st["bf_items_receiver"], st["bf_items_source"]]}},
        "roofline": {"bound": "valu-fp32", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_band<8,4> (receiver side)", "flops_per_cmac": 8,
                     "avg_launch_ms": band_ms},
        "cpu_baseline": cpu,
        "parity": parity,
        "breakdown_ms": {"forward": float(np.mean(fwd)), "tables": float(np.mean(tab)),
                         "receiver": rcv_ms, "source": float(np.mean(src)),
                         "butterfly_receiver": float(np.mean(bfr)),
                         "butterfly_source": float(np.mean(bfs)),
                         "inverse": float(np.mean(inv)), "events_total": float(np.mean(rec))},
        "upload_s": upload_s,
        "out_abs_max": float(np.abs(out).max()),
    }
    print(json.dumps(line), flush=True)


def bench_layers(args, cores):
    """Five depth layers of the salt model as five problems in one
    graph-captured step (their preparation plus one migrate launch)."""
    import numpy as np

    from kirch_datuming_amd import configs, kmig
    t0 = time.time()
    log(f"building the five-layer workload on {cores} host threads")
    ws = configs.layers(threads=cores, overlap=args.overlap)
    total = sum(w.updates for w in ws)
    log(f"workload ready in {time.time() - t0:.1f} s: {total:.4e} updates in 5 layers")
    ms = []
    tu = time.perf_counter()
    for w in ws:
        kv = dict(w.kv, precision=args.precision)
        ax = w.data_axes
        p = kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])
        m = kmig.Migrator(p)
        m.set_tables(0, w.stab, w.sder, w.s_axes)
        m.set_tables(1, w.rtab, w.rder, w.r_axes)
        m.set_data(w.data, ax)
        ms.append(m)
    upload_s = time.perf_counter() - tu
    b = kmig.Batch(ms)
    for i in range(args.warmup):
        b.run()
        b.sync()
        log(f"warmup {i + 1}/{args.warmup}: {b.stats()['ms']:.1f} ms")
    step_ms = []
    tstart = time.perf_counter()
    for i in range(args.steps):
        b.run()
        b.sync()
        step_ms.append(b.stats()["ms"])
    wall = time.perf_counter() - tstart
    st = b.stats()
    value = total * args.steps / wall
    log(f"timed {args.steps} steps in {wall:.3f} s ({st['graph_nodes']} graph nodes)")
    cigs = [m.cig() for m in ms]
    stitch = None
    if args.overlap > 0:
        # the layers overlap: blend them into the full ADCIG on the GPU
        # (layercom_cig's default 15-sample cosine blend); the wall time
        # includes the device scratch and the copy of the volume to the host
        start, end = [w.z_range[0] for w in ws], [w.z_range[1] for w in ws]
        b.stitch(start, end)
        ts = time.perf_counter()
        adcig = b.stitch(start, end)
        stitch = {"ms_with_copy_out": (time.perf_counter() - ts) * 1e3, "blend": 15,
                  "ranges": [list(r) for r in zip(start, end)]}
        log(f"stitched {adcig.shape} in {stitch['ms_with_copy_out']:.1f} ms (with copy-out)")
    else:
        # C5's layers abut (nz 80/70/90/100/60 = 400): the gathers concatenate
        adcig = configs.stack_layers(cigs)
    cpu = parity = None
    if not args.no_cpu and args.cpu_shots > 0:
        from oracle import oracle as O
        subs = [w.shot_subset([w.data.shape[0] // 2]) for w in ws]
        upd = sum(sw.updates for sw in subs)
        log("CPU oracle (float) on the middle shot of each layer")
        tc = time.perf_counter()
        cc = [O.migrate(w.kv, *sw.args(), precision="f32")[0] for w, sw in zip(ws, subs)]
        tcpu = time.perf_counter() - tc
        c64 = [O.migrate(w.kv, *sw.args(), precision="f64")[0] for w, sw in zip(ws, subs)]
        g = [kmig.migrate(dict(w.kv, precision=args.precision), *sw.args())[0]
             for w, sw in zip(ws, subs)]
        cpu = {"value": upd / tcpu, "unit": "updates/s", "cores": cores, "kind": "port",
               "sample": f"middle shot of each of the 5 layers ({upd:.3e} updates), {tcpu:.1f} s"}
        parity = {"rel_l2_gpu_vs_f64_per_layer": [O.rel_l2(a, r) for a, r in zip(g, c64)],
                  "rel_l2_cpu_f32_vs_f64_per_layer": [O.rel_l2(a, r) for a, r in zip(cc, c64)]}
    out = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else f"{args.precision} tables, f32 math",
        "data": "synthetic",
        "config": {"workload": "five salt layers nz 80/70/90/100/60 x 1000"
                               + (f" (layers 2-5 extended {args.overlap} samples up)"
                                  if args.overlap else "") +
                               ", shots 768/640/512/384/256, receivers 401/361/321/281/241, "
                               "nt=2001, one graph-captured step", "config": "layers",
                   "updates_per_step": total, "precision": args.precision,
                   "graph_nodes": st["graph_nodes"]},
        "roofline": {"bound": "hbm", "achieved": total * BYTES_PER_UPDATE[args.precision] /
                     (float(np.mean(step_ms)) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": total * BYTES_PER_UPDATE[args.precision] /
                     (float(np.mean(step_ms)) * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_migrate_multi (whole graph step timed: a lower bound)",
                     "algorithmic_bytes_per_update": BYTES_PER_UPDATE[args.precision]},
        "cpu_baseline": cpu, "parity": parity, "upload_s": upload_s,
        "adcig_shape": list(adcig.shape), "adcig_abs_max": float(np.abs(adcig).max()),
        "stitch": stitch,
    }
    b.close()
    for m in ms:
        m.close()
    print(json.dumps(out), flush=True)


def bench_surrogate(args, cores):
    """Learned-surrogate form of the small workload: the network is pretrained
    on fast-marching tables of seeded random velocity models, fitted to the
    workload's own tables (and, for comparison, fitted from scratch), then
    evaluated inside every step's tuple preparation on both sides, with the
    tables as the per-point fallback (|dT| > 1 ms) and the judge."""
    import numpy as np

    from kirch_datuming_amd import configs, kmig, nn
    from oracle import oracle as O
    FLOP_PER_EVAL = 2.0 * (128 * 48 + 3 * 128 * 128 + 4 * 128)
    BF16_PEAK = 2.5e15
    t0 = time.time()
    w = configs.small(threads=cores)
    ns, nh = w.data.shape[0], w.data.shape[1]
    pos = np.concatenate([w.s_axes[2][1] + np.arange(ns) * w.s_axes[2][2],
                          w.r_axes[2][1] + np.arange(nh) * w.r_axes[2][2]])
    rec_t = configs.table_records(np.concatenate([w.stab, w.rtab]),
                                  np.concatenate([w.sder, w.rder]), pos, 8192, 20260606)
    log(f"small workload and {rec_t.shape[0]} table records in {time.time() - t0:.1f} s")
    t0 = time.time()
    rec_g = configs.grf_records(nmodels=args.grf_models, threads=cores) \
        if args.grf_models > 0 and not args.nn_weights else None
    t_grf = time.time() - t0
    if rec_g is not None:
        log(f"{rec_g.shape[0]} records from {args.grf_models} random models in {t_grf:.1f} s")
    Z, X = np.meshgrid(w.s_axes[0][1] + np.arange(w.s_axes[0][0]) * w.s_axes[0][2],
                       w.s_axes[1][1] + np.arange(w.s_axes[1][0]) * w.s_axes[1][2])

    def check(net):
        err = []
        for k in range(ns):
            s = pos[k]
            o = net.eval(np.stack([np.full(Z.size, s), Z.ravel(), X.ravel()], 1).astype(np.float32))
            err.append(np.abs(o[:, 0] - w.stab[k].ravel()))
        e = np.concatenate(err)
        return {"mean_abs_dt_s": float(e.mean()), "p99_abs_dt_s": float(np.quantile(e, 0.99)),
                "max_abs_dt_s": float(e.max()), "frac_within_1ms": float((e <= 1e-3).mean())}

    nets, report = {}, {}
    if args.nn_weights:
        net = nn.Surrogate(**configs.SURROGATE)
        net.params = np.load(args.nn_weights)
        nets["loaded"] = net
        report["loaded"] = {"file": os.path.basename(args.nn_weights), "check": check(net)}
    if not args.nn_weights and rec_g is not None and args.pretrain_steps > 0:
        pre = nn.Surrogate(**configs.SURROGATE)
        pre.set_data(rec_g)
        lp = pre.train(args.pretrain_steps)
        report["pretrain"] = {"steps": args.pretrain_steps, "loss_first": float(lp[0]),
                              "loss_last": float(lp[-1]), **pre.stats(),
                              "check_vs_workload_tables": check(pre)}
        pre.set_data(rec_t)
        lf = pre.train(args.fit_steps)
        nets["pretrained_then_fitted"] = pre
        report["pretrained_then_fitted"] = {"steps": args.fit_steps,
                                            "loss_last": float(lf[-1]) if len(lf) else None,
                                            **pre.stats(), "check": check(pre)}
    if not args.nn_weights:
        scratch = nn.Surrogate(**configs.SURROGATE)
        scratch.set_data(rec_t)
        ls = scratch.train(args.fit_steps)
        nets["fitted_from_scratch"] = scratch
        report["fitted_from_scratch"] = {"steps": args.fit_steps,
                                         "loss_last": float(ls[-1]) if len(ls) else None,
                                         **scratch.stats(), "check": check(scratch)}
    best = min(nets, key=lambda k: report[k]["check"]["mean_abs_dt_s"])
    net = nets[best]
    if args.save_nn_weights:
        np.save(args.save_nn_weights, net.params)
    log(f"surrogate used: {best}; check {report[best]['check']}")

    kv = dict(w.kv, variant=args.variant)
    if args.tile:
        kv["tile"] = args.tile
    ax = w.data_axes
    p = kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])

    def migrator(surrogate, data_w, per_slot=True):
        m = kmig.Migrator(p)
        m.set_tables(0, w.stab, w.sder, w.s_axes)
        m.set_tables(1, w.rtab, w.rder, w.r_axes)
        m.set_data(data_w.data, data_w.data_axes)
        if surrogate:
            for side in (0, 1):
                m.set_surrogate(side, net, 1e-3)
                m.set_surrogate_fallback(side, per_slot)
        return m

    def timed(m):
        for _ in range(args.warmup):
            m.run()
            m.sync()
        tup, tot = [], []
        ts = time.perf_counter()
        for _ in range(args.steps):
            m.run()
            m.sync()
            st = m.stats()
            tup.append(st["ms_tuple_prep"])
            tot.append(st["ms_total"])
        wall = time.perf_counter() - ts
        return wall, float(np.mean(tup)), float(np.mean(tot))

    mt = migrator(False, w)
    wall_t, tup_t, tot_t = timed(mt)
    mt.close()
    # per-point fallback: the network wherever a point is within 1 ms
    mq = migrator(True, w, per_slot=False)
    wall_q, tup_q, tot_q = timed(mq)
    sst_q = [mq.surrogate_stats(k) for k in (0, 1)]
    mq.close()
    # per-slot fallback (the default): a slot takes the network only where
    # its worst point is within 1 ms
    m = migrator(True, w)
    wall, tup_n, tot_n = timed(m)
    sst = [m.surrogate_stats(k) for k in (0, 1)]
    m.close()
    points = sum(s["points"] for s in sst)
    used = sum(s["used"] for s in sst)
    nn_ms = max(tup_n - tup_t, 1e-6)
    mfma = points * FLOP_PER_EVAL / (nn_ms * 1e-3)
    value = w.updates * args.steps / wall
    log(f"surrogate step {wall / args.steps * 1e3:.2f} ms ({value:.3e} updates/s), table step "
        f"{wall_t / args.steps * 1e3:.2f} ms; network step ~{nn_ms:.2f} ms, "
        f"{mfma / 1e12:.1f} TFLOP/s; {used / points:.3f} of the points from the network")

    # parity on evenly spaced shots: surrogate and table CIGs against FP64
    k = max(2, args.cpu_shots)
    shots = [ns * (j + 1) // (k + 1) for j in range(k)]
    step = shots[1] - shots[0]
    shots = [shots[0] + j * step for j in range(k)]
    subw = w.shot_subset(shots)
    cig = {}
    for name, sur, slot in (("table", False, True), ("surrogate", True, True),
                            ("surrogate_per_point", True, False)):
        ms = migrator(sur, subw, per_slot=slot)
        ms.run()
        ms.sync()
        cig[name] = ms.cig()
        ms.close()
    cig64 = O.migrate(w.kv, *subw.args(), precision="f64")[0]
    cpu = None
    if not args.no_cpu:
        tc = time.perf_counter()
        O.migrate(w.kv, *subw.args(), precision="f32")
        tcpu = time.perf_counter() - tc
        cpu = {"value": subw.updates / tcpu, "unit": "updates/s", "cores": cores, "kind": "port",
               "sample": f"shots {shots} ({subw.updates:.3e} updates), {tcpu:.1f} s, "
                         f"tables (the reference has no surrogate)"}
    table_bytes = 4 * 2 * (w.stab.size + w.rtab.size)
    out = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 migration, bf16 network (fp32 accumulate)", "data": "synthetic",
        "config": {"workload": w.name + ", traveltimes from the learned surrogate (table "
                               "fallback where |dT| > 1 ms)", "config": "surrogate",
                   "updates_per_step": w.updates, "network": "48 Fourier -> 4x128 SiLU -> 4"},
        "roofline": {"bound": "mfma", "achieved": mfma / 1e12, "peak": BF16_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": mfma / BF16_PEAK, "traffic": None,
                     "kernel": "k_nn_infer (network step = tuple preparation with the network "
                               "minus without it)",
                     "flop_per_evaluation": FLOP_PER_EVAL, "evaluations_per_step": points},
        "cpu_baseline": cpu,
        "table_step": {"ms_per_step": wall_t / args.steps * 1e3,
                       "updates_per_s": w.updates * args.steps / wall_t,
                       "tuple_prep_ms": tup_t, "events_total_ms": tot_t},
        "surrogate_step": {"fallback": "per slot (shot or receiver position)",
                           "tuple_prep_ms": tup_n, "events_total_ms": tot_n,
                           "points_from_network": used, "points": points,
                           "max_residual_s": max(s["max_residual"] for s in sst)},
        "surrogate_per_point_step": {
            "fallback": "per point", "ms_per_step": wall_q / args.steps * 1e3,
            "tuple_prep_ms": tup_q, "events_total_ms": tot_q,
            "points_from_network": sum(s["used"] for s in sst_q), "points": points},
        "parity": {"sample": f"shots {shots}",
                   "rel_l2_surrogate_vs_f64": O.rel_l2(cig["surrogate"], cig64),
                   "rel_l2_surrogate_per_point_vs_f64": O.rel_l2(cig["surrogate_per_point"],
                                                                  cig64),
                   "rel_l2_table_vs_f64": O.rel_l2(cig["table"], cig64)},
        "training": report, "network_used": best, "grf_records_s": t_grf,
        "memory_bytes": {"tables": table_bytes, "network_fp32": 4 * nn.PARAMS},
    }
    for q in nets.values():
        q.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
