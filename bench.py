"""VELO reconstruction throughput on one MI355X.

One step = the whole VELO path (super-pixel decoding + masked clustering, phi sort,
search_by_triplet, three-hit filter, consolidation, simplified Kalman fit) over one slice
of 1000 synthetic events with the raw banks already resident in HBM. Slices are dealt to
ten HIP streams in turn (the process runs with twelve hardware queues), one sequence per
stream, which is how Allen measures GPU throughput (one stream per host thread, scanned
over the stream count); the one-stream figure is reported beside it.

Prints ONE JSON line (rank 0). Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # One slice of 1000 events; nu ~ Poisson(7.6) collisions x NegBinomial(16, k=1.5) tracks.
    "pileup_1000": dict(n_events=1000, pileup=1, nu=7.6, seed=1),
    # 200 events of Poisson(120) tracks from one PV (the reference's CPU-runnable case).
    "single_pv_200": dict(n_events=200, pileup=0, mean_tracks=120.0, seed=1),
    # 20k events of one PV with Poisson(120) tracks, 5 % of them nu = 20 pile-up events.
    "batched_20k": dict(n_events=20000, pileup=0, mean_tracks=120.0, high_occupancy_fraction=0.05, high_nu=20.0,
                        seed=1),
    # 10k events at nu ~ Poisson(15), 5 % noise pixels, 1 % dead ASICs.
    "stress_10k": dict(n_events=10000, pileup=1, nu=15.0, noise_fraction=0.05, dead_asic_fraction=0.01, seed=1),
}

HBM_PEAK_GBS = 8000.0                    # MI355X HBM3E
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz, non-packed lane ops/s
VALU_ISSUE_PEAK_G = 1024 * 2.4e9 / 4 / 1e9  # 1024 SIMDs, one wave64 VALU instruction per 4 cycles each
LDS_PEAK_GBS = 256 * 256 * 2.4e9 / 1e9    # 256 CUs x 256 B/clk (CDNA4, 128-bit reads) x 2.4 GHz
CPU_SHARE = 16                            # host cores that belong to one GPU of the box


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload: str, stage: str):
    """HBM bytes per launch of `stage`'s kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    committed for this workload (the last column of the stage's row), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) >= 3 and cols[0] == stage and cols[-1].isdigit():
            return int(cols[-1])
    return None


def measured_valu_instructions(workload: str):
    """Wave-level VALU instructions per launch of search_by_triplet from the rocprofv3 shader-counter
    summary committed for this workload (the first column of its SQ_INSTS_VALU row), or None."""
    path = os.path.join(ROOT, "profiles", f"search_by_triplet_{workload}_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) >= 2 and cols[0] == "SQ_INSTS_VALU" and cols[1].isdigit():
            return int(cols[1])
    return None


def physics_of(seq, ev) -> dict:
    """Efficiency, ghost and clone rates of the last run's tracks against the generator's truth (01_velo)."""
    from allen_amd import checker
    n = ev.n_events
    off_all = seq.fetch("offsets_all_velo_tracks", np.uint32, n + 1)
    n_tracks = int(off_all[-1])
    hits_n = seq.fetch("offsets_velo_track_hit_number", np.uint32, n_tracks + 1)
    track_hits = seq.fetch("velo_track_hits", np.uint8, 4 * int(hits_n[-1]))
    return checker.check_buffers(ev, off_all, hits_n, track_hits).as_dict()


def algorithmic_bytes(ev, counts: dict) -> dict:
    """Algorithmic HBM bytes of each stage for one step: clustering reads the raw banks, the size and
    type arrays and the offsets, and writes 18 B per cluster and 4 B per module pair; the search reads
    18 B per cluster and the module-pair offsets and counts, and writes a used flag per cluster, the
    track hits, 6 B per three-hit track, 10 B per seed (tracklet and ring entry) and 16 B of counters
    per event."""
    n = ev.n_events
    sizes16 = ev.sizes.view("<u2")
    n_banks = int(ev.sizes[n]) - int(ev.sizes[0])
    raw_read = int(ev.raw.size) + 2 * n_banks + n_banks + 3 * 4 * (n + 1)
    n_c = counts["clusters"]
    cl = raw_read + 18 * n_c + 4 * 26 * n
    search_read = 18 * n_c + 2 * 4 * 26 * n
    search_write = n_c + counts["track_hit_bytes"] + 6 * counts["three_hit_raw"] + 10 * counts["seeds"] + 16 * n
    return {"clustering": cl, "search": search_read + search_write}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pileup_1000", choices=sorted(WORKLOADS))
    ap.add_argument("--decoding", choices=["super_pixel", "retina"], default="super_pixel",
                    help="raw banks as super pixels (masked clustering) or as retina clusters made by the "
                         "front-end (decode_retinaclusters), Allen's two VELO decodings")
    ap.add_argument("--pv-finder", action="store_true",
                    help="also run the beamline primary-vertex finder on the beamline states (the path's consumer; "
                         "the cpu baseline then runs it too)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline; 0: the host cores this process may use, at most the 16 "
                         "that are one GPU's share of the box (both figures are reported)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU work of the all-core baseline sample (the single-thread sample takes half)")
    ap.add_argument("--search-block", type=int, default=0)
    ap.add_argument("--search-lds-hits", type=int, default=0)
    ap.add_argument("--search-ring-window", type=int, default=0,
                    help="LDS window of the ring of tracks to follow (power of two, 64-512; 0: 512)")
    ap.add_argument("--search-list-cap", type=int, default=0,
                    help="capacity of the seeding's LDS list of unused hits (multiple of 16, up to 1024; 0: 1024)")
    ap.add_argument("--search-event-order", choices=["largest_first", "natural"], default="largest_first")
    ap.add_argument("--search-split", type=int, default=0,
                    help="the largest 1/N of a slice's events run in a second search launch, on a second stream of "
                         "the sequence, in 256-lane workgroups with 4096 LDS hits; 0: one launch (the default: the "
                         "split shortens one slice's search but the extra stream per slice costs more throughput "
                         "than it gains when several slices are in flight)")
    ap.add_argument("--streams", type=int, default=10,
                    help="slices in flight: one sequence and one HIP stream per slice, as Allen runs one stream per "
                         "host thread and measures GPU throughput over the stream count; steps are dealt to the "
                         "streams in turn (10 streams on 12 hardware queues measured best on one MI355X, within a "
                         "plateau from 10 to 14 streams on 11 to 16 queues; 3 with HIP's default of 4)")
    ap.add_argument("--hw-queues", type=int, default=12,
                    help="hardware queues of this process (HIP's GPU_MAX_HW_QUEUES, set before the runtime starts; "
                         "HIP's default is 4): one per stream in flight plus the stream that times them")
    ap.add_argument("--persistent", action="store_true",
                    help="search_by_triplet as a persistent grid whose workgroups take events from a work counter")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="before the timed steps, time whole steps for each search_by_triplet workgroup size and "
                         "LDS hit budget of a small grid, in this run's own streams and workload, and keep the "
                         "fastest (outputs are checked to be unchanged)")
    ap.add_argument("--graph", action="store_true",
                    help="time replays of the whole sequence captured as one HIP graph instead of eager runs")
    ap.add_argument("--nn-pruned", action=argparse.BooleanOptionalAction, default=True,
                    help="after the plain path, time the NN-pruned path (learned seed scorer) on the same events and "
                         "streams and report it beside the plain one, as the metric names both")
    ap.add_argument("--scorer", default="shipped",
                    help="scorer of the NN-pruned leg: 'shipped' (the package's trained scorer) or an .npz of "
                         "parameters (allen_amd.scorer)")
    ap.add_argument("--threshold", type=float, default=None,
                    help="logit below which a seed is pruned (default: the shipped scorer's operating threshold)")
    ap.add_argument("--microbench", action="store_true",
                    help="only measure this device's vector-ALU FMA issue rate, LDS read bandwidth and HBM copy "
                         "bandwidth (velo_microbench) and print them beside the published peaks")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if not 1 <= args.hw_queues <= 32:
        print(f"bench.py: --hw-queues {args.hw_queues} is outside 1..32", file=sys.stderr)
        sys.exit(2)
    # Read by the HIP runtime when it starts, which is at the first GPU call below.
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the VELO path shards by events "
              f"over independent GPUs, and this program measures one.", file=sys.stderr)
        sys.exit(2)

    import torch
    from allen_amd import synth, velo

    assert torch.cuda.is_available(), "no HIP device"
    if args.microbench:
        valu = velo.microbench("valu_lane_ops_per_s")
        lds = velo.microbench("lds_bytes_per_s")
        hbm = velo.microbench("hbm_copy_bytes_per_s")
        # The HIP runtime's own device copy of the same size, timed the same way (five copies
        # after one untimed copy, HIP events).
        src = torch.ones((1 << 30) // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(5):
            dst.copy_(src)
        t1.record()
        torch.cuda.synchronize()
        hbm_rt = 5 * 2 * (1 << 30) / (t0.elapsed_time(t1) / 1e3)
        del src, dst
        print(json.dumps({"microbench": {
            "valu": {"measured": round(valu / 1e12, 2), "published": round(VALU_PEAK_TOPS, 2), "unit": "T lane-ops/s",
                     "frac": round(valu / 1e12 / VALU_PEAK_TOPS, 3)},
            "lds": {"measured": round(lds / 1e9, 1), "published": round(LDS_PEAK_GBS, 1), "unit": "GB/s",
                    "frac": round(lds / 1e9 / LDS_PEAK_GBS, 3)},
            "hbm_copy": {"measured": round(hbm / 1e9, 1), "published": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(hbm / 1e9 / HBM_PEAK_GBS, 3)},
            "hbm_copy_runtime": {"measured": round(hbm_rt / 1e9, 1), "published": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(hbm_rt / 1e9 / HBM_PEAK_GBS, 3)}}}), flush=True)
        return
    if args.copy_calibration:
        n_bytes = args.copy_calibration << 20
        src = torch.ones(n_bytes // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n_bytes, "launches": 5}), flush=True)
        return
    wl = dict(WORKLOADS[args.workload])
    if args.decoding == "retina":
        wl["retina"] = 1
    t0 = time.time()
    ev = synth.generate(**wl, keep_truth=True, n_threads=min(16, os.cpu_count() or 1))
    log(f"[bench] generated {ev.n_events} events ({ev.raw.size / 1e6:.2f} MB raw) in {time.time() - t0:.1f}s")
    geometry, beamline = synth.read_geometry(), synth.read_beamline()
    constants = velo.VeloConstants(geometry, beamline)
    dev = velo.DeviceEvents(ev)
    knobs = {"search_largest_first": int(args.search_event_order == "largest_first"),
             "search_persistent": int(args.persistent), "search_split": args.search_split}
    if args.search_block:
        knobs["search_block"] = args.search_block
    if args.search_lds_hits:
        knobs["search_lds_hits"] = args.search_lds_hits
    if args.search_ring_window:
        knobs["search_ring_window"] = args.search_ring_window
    if args.search_list_cap:
        knobs["search_list_cap"] = args.search_list_cap
    # One sequence per stream (slices in flight), configured alike; the first is also the one the
    # diagnostics below run on.
    n_streams = max(1, args.streams)
    seqs = [velo.VeloSequence(constants, ev.n_events, ev.raw.size) for _ in range(n_streams)]
    seq = seqs[0]
    for sq in seqs:
        for k, v in knobs.items():
            sq.set_tuning(k, v)
        if args.pv_finder:
            sq.set_tuning("pv_finder", 1)
    stream = torch.cuda.current_stream()
    streams = [stream] if n_streams == 1 and not args.graph else [torch.cuda.Stream() for _ in range(n_streams)]
    # Set-up: every sequence runs once on its own stream. The HIP runtime makes a stream's hardware
    # queue at the stream's first launch (milliseconds), which must not fall into the timed steps.
    for sq, st in zip(seqs, streams):
        sq.run(dev, stream=st)
    torch.cuda.synchronize()

    host = {"ms": 0.0}  # host time of the last timed() call's launch loop

    def timed(steps: int, warmup: int, stage: str | None, used: int = n_streams) -> float:
        """Milliseconds of `steps` steps dealt to the first `used` streams in turn (eager runs or
        graph replays), HIP events around them on the current stream; `stage` is bracketed in the
        first sequence."""
        if args.graph:
            for sq, st in zip(seqs[:used], streams[:used]):
                sq.time_stage(None)
                sq.capture(dev, st)

        def launch(i):
            sq, st = seqs[i % used], streams[i % used]
            if args.graph:
                sq.replay(st)
            else:
                sq.run(dev, stream=st)

        for i in range(warmup):
            launch(i)
        torch.cuda.synchronize()
        if stage and not args.graph:
            seq.time_stage(stage)
        main_stream = torch.cuda.current_stream()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(main_stream)
        for st in streams[:used]:
            st.wait_event(start)
        t_host = time.perf_counter()
        for i in range(steps):
            launch(i)
        host["ms"] = (time.perf_counter() - t_host) * 1e3
        for st in streams[:used]:
            main_stream.wait_stream(st)
        end.record(main_stream)
        torch.cuda.synchronize()
        return start.elapsed_time(end)

    tuned = None
    tuning = None
    if args.autotune:
        # Workgroup size and LDS hit budget of search_by_triplet, chosen by timing whole steps in
        # this run's own regime (streams, graph, workload); outputs must not change. The scored
        # search keeps its fixed workgroup of 256.
        from allen_amd import autotune
        t0 = time.time()
        res = autotune.tune_throughput(seqs, lambda: timed(20, 3, None) / 20, ev.n_events, log=log)
        tuned = res.knobs
        knobs.update(tuned)
        tuning = {"default_ms_per_step": round(res.default_search_ms, 4), "best_ms_per_step": round(res.search_ms, 4),
                  "trials": res.trials, "seconds": round(time.time() - t0, 1)}
        log(f"[bench] autotune ({time.time() - t0:.1f}s, {len(res.trials)} trials): {res.knobs} "
            f"{res.search_ms:.4f} ms per step (block 256: {res.default_search_ms:.4f} ms)")

    for _ in range(args.warmup):
        seq.run(dev)
    torch.cuda.synchronize()
    status = seq.status()
    if status & velo.STATUS_OVERFLOW_MASK:
        log(f"[bench] WARNING status {status:#x}: a fixed-capacity structure overflowed")

    # Per-stage times (diagnostic, outside the timed region) to find the dominant kernel.
    stage_ms = {}
    for stage in ["candidates", "estimate", "clustering", "search", "three_hit_filter", "consolidate", "kalman"] + \
            (["pv"] if args.pv_finder else []):
        seq.time_stage(stage)
        for _ in range(2):
            seq.run(dev)
        ms, runs = seq.stage_ms()
        stage_ms[stage] = ms / runs
    dominant = max(stage_ms, key=stage_ms.get)
    log(f"[bench] stage ms: " + ", ".join(f"{k}={v:.3f}" for k, v in stage_ms.items()) + f" -> {dominant}")

    # Work counts and output sizes of one step.
    seq.time_stage(None)
    seq.set_tuning("count_work", 1)
    seq.run(dev)
    work = seq.fetch("search_work", np.uint64, 4)
    seq.set_tuning("count_work", 0)
    n = ev.n_events
    # Per-event workgroup durations of search_by_triplet (counting run, wall clock).
    tick_us = 1e3 / velo.wall_clock_khz()
    clk = seq.fetch("search_event_clock", np.uint64, 10 * n).reshape(n, 10).astype(np.int64)
    us = (clk[:, 1] - clk[:, 0]) * tick_us
    span_us = (clk[:, 1].max() - clk[:, 0].min()) * tick_us
    last_start_us = (clk[:, 0].max() - clk[:, 0].min()) * tick_us
    event_hits = np.diff(seq.fetch("offsets_estimated_input_size", np.uint32, 26 * n + 1)[::26].astype(np.int64))
    slope, intercept = (float(v) for v in np.polyfit(event_hits, us, 1)) if n > 1 else (0.0, 0.0)
    big = int(np.argmax(event_hits))
    med = int(np.argsort(event_hits)[n // 2])
    phases = ["forwarding", "compaction", "seed1_search", "seed1_pendulum", "seed1_wait", "seed_step2", "seed_step3",
              "other"]
    search_events = {"p50_us": round(float(np.median(us)), 2), "p90_us": round(float(np.percentile(us, 90)), 2),
                     "max_us": round(float(us.max()), 2), "span_us": round(float(span_us), 2),
                     "last_start_us": round(float(last_start_us), 2), "us_per_hit": round(slope, 5),
                     "us_fixed": round(intercept, 2), "largest_event_hits": int(event_hits[big]),
                     "largest_event_us": round(float(us[big]), 2),
                     "largest_event_phase_us": {k: round(float(clk[big, 2 + i] * tick_us), 2)
                                                for i, k in enumerate(phases)},
                     "median_event_hits": int(event_hits[med]),
                     "median_event_phase_us": {k: round(float(clk[med, 2 + i] * tick_us), 2)
                                               for i, k in enumerate(phases)}}
    log(f"[bench] search workgroups (counting run): {search_events}")
    mcn = seq.fetch("module_cluster_num", np.uint32, 26 * n)
    off_all = seq.fetch("offsets_all_velo_tracks", np.uint32, n + 1)
    atomics = seq.fetch("atomics_velo", np.uint32, 3 * n)
    n_tracks = int(off_all[-1])
    hit_n = seq.fetch("velo_track_hit_number", np.uint32, n_tracks)
    counts = {
        "clusters": int(mcn.sum()),
        "tracks": n_tracks,
        "track_hit_bytes": int((2 * hit_n + 1).sum()),
        "three_hit_raw": int(atomics[0::3].sum()),
        "seeds": int(atomics[1::3].sum()),
    }
    log(f"[bench] {counts}")
    # The counting run above uses the plain search; the physics is that of the timed path.
    seq.run(dev)
    physics = physics_of(seq, ev)
    log(f"[bench] physics: {physics}")

    # The same steps on one stream, for reference (outside the timed region).
    single_ms = timed(args.steps, args.warmup, None, 1) / args.steps if n_streams > 1 else None
    # Timed region: K steps; without a graph the dominant stage is bracketed by HIP events on its
    # stream in the first sequence.
    total_ms = timed(args.steps, args.warmup, dominant)
    host_ms_per_step = host["ms"] / args.steps
    if args.graph:
        # The graph carries no stage events; the dominant kernel's launch time is the eager one.
        kern_avg_ms = stage_ms[dominant]
    else:
        kern_ms, kern_runs = seq.stage_ms()
        kern_avg_ms = kern_ms / kern_runs
        seq.time_stage(None)
    for sq in seqs[1:min(n_streams, args.steps)]:  # the sequences the timed steps ran
        if physics_of(sq, ev) != physics:
            raise RuntimeError("a second stream's slice gave other tracks than the first")
    ms_per_step = total_ms / args.steps
    events_per_s = n * args.steps / (total_ms / 1e3)
    status = seq.status()

    alg = algorithmic_bytes(ev, counts)
    dom_bytes = alg.get(dominant)
    if dom_bytes is None:
        dom_bytes = alg["clustering"]
    # The dominant kernel's launch time is its one-stream time (HIP events around the stage on the
    # run's stream, nothing else on the GPU), which the kernel trace of the same command reproduces;
    # in the multi-stream timed steps its launches share the CUs with other slices' kernels.
    launch_ms = stage_ms[dominant]
    achieved = dom_bytes / (launch_ms / 1e3) / 1e9
    traffic_name = args.workload + ("_retina" if args.decoding == "retina" else "")
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(traffic_name, dominant),
                "algorithmic_bytes": dom_bytes, "avg_launch_ms": round(launch_ms, 4),
                "launch_ms_in_timed_steps": round(kern_avg_ms, 4)}
    # search_by_triplet's own bounds are the vector ALUs and LDS, not HBM: per scored candidate
    # (seeding pair or forwarding h2) 9 VALU lane-ops and 15 B of hit data from LDS (x, y, z, phi,
    # flags); per extrapolation ~12 lane-ops (hit_phi_16 with its division); per binary-search probe
    # 3 lane-ops and a 2 B phi from LDS.
    evals = int(work[0]) + int(work[1])
    ops = 9 * evals + 12 * int(work[2]) + 3 * int(work[3])
    lds_bytes = 15 * evals + 2 * int(work[3])
    search_s = stage_ms["search"] / 1e3
    search_tops = ops / search_s / 1e12
    search_lds = lds_bytes / search_s / 1e9
    search_alu_lds = {"valu": {"ops_per_launch": ops, "achieved": round(search_tops, 4),
                               "peak": round(VALU_PEAK_TOPS, 2), "unit": "T lane-ops/s",
                               "frac": round(search_tops / VALU_PEAK_TOPS, 5)},
                      "lds": {"bytes_per_launch": lds_bytes, "achieved": round(search_lds, 2),
                              "peak": round(LDS_PEAK_GBS, 1), "unit": "GB/s",
                              "frac": round(search_lds / LDS_PEAK_GBS, 6)},
                      "hbm": {"achieved": round(alg["search"] / search_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(alg["search"] / search_s / 1e9 / HBM_PEAK_GBS, 5)},
                      "avg_launch_ms": round(stage_ms["search"], 4),
                      "work": {"seed_evaluations": int(work[0]), "forward_evaluations": int(work[1]),
                               "extrapolations": int(work[2]), "search_probes": int(work[3])}}
    # The instructions the kernel actually issues (the committed shader counters of this workload,
    # at the line's default knobs) against the chip's VALU issue rate: one launch alone on the GPU.
    valu_instr = measured_valu_instructions(traffic_name) if args.decoding == "super_pixel" else None
    if valu_instr:
        issue_g = valu_instr / search_s / 1e9
        search_alu_lds["valu_issue"] = {"instructions_per_launch": valu_instr, "achieved": round(issue_g, 2),
                                        "peak": round(VALU_ISSUE_PEAK_G, 1), "unit": "G wave-instructions/s",
                                        "frac": round(issue_g / VALU_ISSUE_PEAK_G, 4)}
    if dominant == "search":
        roofline["search_alu_lds"] = search_alu_lds
    # Decoding and clustering against HBM: the algorithmic bytes of K1 + K2 + K3 over the sum of
    # their one-stream launch times.
    cl_ms = stage_ms["candidates"] + stage_ms["estimate"] + stage_ms["clustering"]
    cl_gbs = alg["clustering"] / (cl_ms / 1e3) / 1e9
    extra = {"clustering_hbm": {"achieved": round(cl_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(cl_gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes": alg["clustering"],
                                "launch_ms": round(cl_ms, 4),
                                "stages": ["candidates", "estimate", "clustering"],
                                "traffic": measured_traffic(traffic_name, "clustering")},
             "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
             "search_events": search_events,
             "physics": physics}
    if dominant != "search":
        extra["search_alu_lds"] = search_alu_lds

    # The NN-pruned leg: the same events, streams and search knobs with the learned seed scorer
    # pruning seeds (its groups of seeds must lie in one wave: workgroups of 64, 128 or 256
    # lanes, else 256), timed the same way.
    nn = None
    if args.nn_pruned:
        from allen_amd import scorer
        if args.scorer == "shipped":
            params, shipped_threshold = scorer.trained()
        else:
            params, shipped_threshold = scorer.load_params(args.scorer), 0.0
        threshold = shipped_threshold if args.threshold is None else args.threshold
        nn_knobs = {k: v for k, v in knobs.items() if k in ("search_block", "search_lds_hits", "search_ring_window",
                                                             "search_list_cap")}
        if nn_knobs.get("search_block", 256) not in (64, 128, 256):
            nn_knobs = {"search_block": 256, "search_lds_hits": 0, "search_ring_window": 0, "search_list_cap": 0}
        for sq in seqs:
            for k, v in nn_knobs.items():
                sq.set_tuning(k, v)
            sq.set_scorer(params, threshold)
        nn_ms = timed(args.steps, args.warmup, None) / args.steps
        seq.run(dev)
        nn_physics = physics_of(seq, ev)
        nn = {"events_per_s": round(n / (nn_ms / 1e3), 1), "ms_per_step": round(nn_ms, 4),
              "speedup_vs_plain": round(ms_per_step / nn_ms, 4), "threshold": threshold,
              "scorer": os.path.basename(args.scorer), "search_knobs": nn_knobs,
              "physics": nn_physics,
              "long_efficiency_loss": round(physics["02_long_efficiency"] - nn_physics["02_long_efficiency"], 6),
              "ghost_rate_change": round(nn_physics["ghost_rate"] - physics["ghost_rate"], 6)}
        log(f"[bench] nn-pruned: {nn}")
        for sq in seqs:
            sq.set_scorer(None)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = args.cpu_threads or min(CPU_SHARE, usable)
        sample = synth.generate(**dict(wl, n_events=min(n, 1000)), keep_truth=False)
        config1 = synth.generate(**WORKLOADS["single_pv_200"], keep_truth=False)

        def cpu_timed(events, n_threads, seconds):
            # One calibration pass, then enough repetitions of the sample for about `seconds` of CPU work.
            _, s1 = oracle.throughput(events, geometry, beamline, n_threads, 1, native=True,
                                      pv_finder=int(args.pv_finder))
            reps = max(1, int(round(seconds / max(s1, 1e-3))))
            eps, secs = oracle.throughput(events, geometry, beamline, n_threads, reps, native=True,
                                          pv_finder=int(args.pv_finder))
            return eps, secs, reps

        eps, secs, reps = cpu_timed(sample, threads, args.cpu_seconds)
        eps1, secs1, reps1 = cpu_timed(sample, 1, args.cpu_seconds / 2)
        c1, c1s, c1r = cpu_timed(config1, threads, args.cpu_seconds / 4)
        c1_1, c1_1s, c1_1r = cpu_timed(config1, 1, args.cpu_seconds / 4)
        log(f"[bench] cpu baseline: {eps:.1f} events/s on {threads} threads, {eps1:.1f} on one; "
            f"single_pv_200: {c1:.1f} / {c1_1:.1f}")
        cpu = {"value": round(eps, 1), "unit": "events/s", "cores": threads, "kind": "port",
               "build": "oracle/Makefile native: g++ -O3 -march=x86-64-v3 (Allen's CPU-target optimisation and CI "
                        "platform level), default contraction",
               "cpu_model": cpu_model(), "host_cores_usable": usable,
               "cores_note": f"{threads} threads: the host cores usable here, at most the {CPU_SHARE} that are one "
                             f"GPU's share of the box" if not args.cpu_threads else "set by --cpu-threads",
               "sample": f"{sample.n_events} events of the same workload x {reps} repetitions, one event per task "
                         f"on {threads} std::threads ({secs:.1f} s); single thread: {sample.n_events} events x "
                         f"{reps1} repetitions ({secs1:.1f} s) -> {eps1:.1f} events/s",
               "single_thread_events_per_s": round(eps1, 1),
               "single_pv_200": {"events_per_s": round(c1, 1), "threads": threads,
                                 "single_thread_events_per_s": round(c1_1, 1),
                                 "sample": f"200 events x {c1r} repetitions ({c1s:.1f} s); single thread x "
                                           f"{c1_1r} ({c1_1s:.1f} s)"}}

    out = {
        "metric": "VELO events/s on 1 MI355X, plain vs NN-pruned, at fixed long-track efficiency",
        "value": round(events_per_s, 1),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, real VELO geometry, v4 super-pixel banks)",
        "config": {"workload": args.workload, "events_per_step": n, "decoding": args.decoding,
                   "pv_finder": bool(args.pv_finder), "path": "plain",
                   "launch": "hip_graph" if args.graph else "eager", "streams": n_streams,
                   "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
                   "search_grid": "persistent" if args.persistent else "one_workgroup_per_event",
                   "search_split": args.search_split,
                   **({"search_knobs": tuned} if tuned else {}),
                   "raw_bytes_per_step": int(ev.raw.size), **{k: v for k, v in wl.items() if k != "n_events"}},
        "status": status,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "long_track_efficiency": physics["02_long_efficiency"],
        **({"nn_pruned": nn} if nn else {}),
        **({"autotune": tuning} if tuning else {}),
        **({"single_stream": {"ms_per_step": round(single_ms, 4), "events_per_s": round(n / (single_ms / 1e3), 1)}}
           if single_ms else {}),
        # Host time to enqueue one step (the launch loop of the timed region): a step cannot
        # complete faster than the host issues it.
        "host_enqueue_ms_per_step": round(host_ms_per_step, 4),
        **extra,
    }
    print(json.dumps(out), flush=True)
    for sq in seqs[1:]:
        sq.close()
    seq.close()
    constants.close()


if __name__ == "__main__":
    main()
