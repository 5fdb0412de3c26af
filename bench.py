#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the plain FP32 substep at ~32M points.

Workload (one MI355X): 4096 x 2048 nodes, h = 1 m, ~31.8M points at 4 per
cell in a seeded random polygonal ice field covering ~95 % of the domain,
Gaussian-random-field wind stress, lognormal per-point cohesion, dt = 1e-4 s.
Inputs are resident in HBM when the timed region starts.  A "step" is one
substep (P2G, node update, G2P with SVD and return map) over all points.

Warm-up runs W eager substeps and then one 16-substep HIP-graph chunk, so the
graph is captured and instantiated before timing.  The timed region replays K
substeps from that graph (16 per replay; a remainder runs eagerly), bracketed
by stream synchronisation; `value` = active points * K / wall time.  A second,
shorter pass of the same substeps runs eagerly with HIP events around each
kernel on the library's stream; it gives the dominant kernel's average launch
time, from which `roofline` prices the kernel's algorithmic bytes per launch
against MI355X's 8 TB/s.

The dominant kernel, k_tile_pipeline, is a G2P2G pass: one launch runs the
node update and G2P of substep n and the P2G of substep n + 1, whose inputs
come from registers.  Its algorithmic bytes per launch are therefore those of
the fused design: per point the G2P reads (utility, cell, pos x2, Fe x4,
Jp_inv, thickness, cohesion: 44 B) and writes (utility, cell, pos x2, vel x2,
Fe x4, Bp x4, Jp_inv: 60 B), 104 B; per node 33 B (P2G writes m, px, py; the
update reads them, the status byte and the 2-float wind).  The two-pass
substep's bytes (168 B per point + 33 B per node) priced on the same launch,
the bytes the pipelined design itself must move (`own_design`) and the HBM
bytes the counters measured (`traffic`) are reported beside them.

Two secondary legs follow unless --no-secondary: the 64 uneven domains
batched against the same domains run one after another (`batched`), and the
64 domains in compression with the learned plastic projection against the
plain forms (`surrogate`).  `cpu_baseline` times the FP64 oracle on the host
cores over a bounded sample of the headline workload.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n_points: int, n_nodes: int, cohesion: bool, complete_frac: float = 1.0,
                      rec_frac: float = 0.0) -> dict:
    """Bytes each kernel must move once per substep (FP32 hssoa, 4 B slots).

    Unfused: P2G reads utility, cell, pos x2, vel x2, Fe x4, Bp x4, Jp_inv,
    thickness (16 slots) and writes m, px, py per node.  The update + G2P
    kernel reads utility, cell, pos x2, Fe x4, Jp_inv, thickness (+ cohesion)
    and writes utility, cell, pos x2, vel x2, Fe x4, Bp x4, Jp_inv (15 slots);
    per node it reads m, px, py, the status byte and the 2-float wind stress.

    Fused (G2P2G, one tile pass per substep, the P2G's point inputs taken
    from registers): the G2P's point bytes plus 33 B per node.

    Pipelined (the bytes that design itself moves): per point the same 10 (+1) slots
    read; position and Fe (6 slots) written every substep, velocity and Bp
    (6 slots) on the substeps that complete a run chunk (complete_frac of
    them), P and Q on recording substeps (rec_frac); Jp_inv, the utility and
    cell words only when they change (not counted).  Per node the node sums
    read and written once (24 B), status (1 B), wind (8 B), cell start (4 B).
    """
    p2g = 64 * n_points + 12 * n_nodes
    g2p = (40 + (4 if cohesion else 0) + 60) * n_points + 21 * n_nodes
    pipe = (40 + (4 if cohesion else 0) + 24 + 24 * complete_frac + 8 * rec_frac) * n_points + 37 * n_nodes
    # G2P2G: the G2P's point reads and writes, the P2G's and the update's node bytes
    fused = (40 + (4 if cohesion else 0) + 60) * n_points + 33 * n_nodes
    return {"p2g": p2g, "g2p": g2p, "substep": p2g + g2p, "pipelined": pipe, "fused": fused}


def measured_traffic(kernel: str, tile: str, grid: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary for this workload and tile (FETCH_SIZE and WRITE_SIZE passes,
    scaled by the same counters over the plain copy kernel), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{kernel}_icefield{grid}_*.md"))):
        for line in open(path):
            m = re.match(r"traffic_bytes_per_launch: tile (\S+) ([0-9.eE+]+)", line.strip())
            if m and m.group(1) == tile:
                return float(m.group(2))
    return None


def cpu_baseline(scene, target_s: float, threads: int, max_steps: int = 50) -> dict:
    """FP64 oracle on the host cores: 1 warm-up substep, then whole substeps
    until about target_s seconds of CPU work (at least 2)."""
    import oracle
    nthr = oracle.threads(threads)
    o = oracle.Oracle.from_scene(scene, 64)
    o.run(0, 1)                       # first-touch / warm-up substep
    first = o.bench(1, 1)
    steps = int(min(max_steps, max(1, round(target_s / max(first, 1e-6)) - 1)))
    secs = first + o.bench(2, steps)
    steps += 1
    n = o.n
    o.close()
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": n * steps / secs, "unit": "particle-updates/s", "cores": nthr, "kind": "port",
            "sample": f"FP64 oracle (OpenMP, {nthr} threads; host nproc {os.cpu_count()}, {model}), {steps} "
                      f"substeps of the same {n}-point workload after 1 warm-up substep, {secs:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--event-steps", type=int, default=48, help="eager substeps timed per kernel with HIP events")
    ap.add_argument("--tile", type=str, default="16x32")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--naive", action="store_true", help="reference launch structure (unsorted, atomics)")
    ap.add_argument("--unfused", action="store_true", help="separate P2G and update + G2P tile kernels")
    ap.add_argument("--det", action="store_true",
                    help="deterministic mode (fixed-point P2G, bit-identical results; unfused kernels)")
    ap.add_argument("--persistent", action="store_true", help="tile passes as persistent launches (work queues)")
    ap.add_argument("--resort-drift", type=int, default=None,
                    help="re-sort a domain when this %% of its points left their sorted cell (0 = off; library default 30)")
    ap.add_argument("--scale", type=float, default=1.0, help="linear domain scale (1.0 = 4096x2048)")
    ap.add_argument("--workload", choices=["icefield", "batch64", "surrogate", "frame", "compressed"],
                    default="icefield",
                    help="icefield: one 4096x2048 domain (the headline); batch64: 64 uneven domains; "
                         "surrogate: the 64 domains under high-strain-rate forcing, plain vs learned projection; "
                         "frame: the per-frame visualisation arrays of the 4096x2048 domain; "
                         "compressed: the plain batched substep of the surrogate workload's domains (fast flow)")
    ap.add_argument("--surrogate-tol", type=float, default=1e-4, help="yield-violation tolerance of the fallback")
    ap.add_argument("--single-handle", action="store_true",
                    help="one handle only (for profiling): the surrogate workload's surrogate leg with eager "
                         "launches; the batch64 workload without its one-handle-per-domain sequential leg")
    ap.add_argument("--copy-only", type=float, default=0.0,
                    help="only time a plain copy of this many GB (HBM reference / counter calibration)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the batched and surrogate legs of the default workload")
    ap.add_argument("--secondary-steps", type=int, default=32, help="timed substeps of each secondary leg")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this single-GPU build "
              "(the multi-GPU slab decomposition is design-only)", file=sys.stderr)
        return 2

    from icempm_2d_sp_amd import abi, scenes
    from icempm_2d_sp_amd.gpu import Simulation

    if args.copy_only > 0:
        import ctypes as C
        nbytes = int(args.copy_only * 1e9)
        out = {"bytes": nbytes}
        for mode, name, moved in [(0, "copy", 2 * nbytes), (1, "read", nbytes), (2, "write", nbytes)]:
            ms = C.c_double(0)
            cal = abi.load_calib()
            rc = cal.ice_stream_bench(0, nbytes, 20, mode, C.byref(ms))
            if rc != 0:
                raise abi.IceError(f"ice_stream_bench failed ({rc}): {cal.ice_calib_last_error().decode()}")
            gbs = moved / (ms.value * 1e-3) / 1e9
            out[name] = {"ms": ms.value, "GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}
        print(json.dumps(out), flush=True)
        return 0

    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)
    if args.workload == "frame":
        return bench_frame(args)
    if args.workload == "compressed":
        return bench_compressed(args)

    gx, gy = int(round(4096 * args.scale)), int(round(2048 * args.scale))
    t0 = time.time()
    scene = scenes.icefield_4096x2048(gx=gx, gy=gy)
    t_gen = time.time() - t0
    print(f"[bench] scene {gx}x{gy}: {scene.n} points, generated in {t_gen:.1f}s", file=sys.stderr, flush=True)

    tw, th = (int(v) for v in args.tile.split("x"))
    flags = scene.params.flags | (abi.FLAG_NAIVE if args.naive else 0) | (abi.FLAG_DETERMINISTIC if args.det else 0)
    pipelined = not (args.naive or args.unfused or args.det)
    opts = {"pipelined": int(pipelined), "persistent": int(args.persistent)}
    if args.resort_drift is not None:
        opts["resort_drift"] = args.resort_drift
    sim = Simulation(scene, flags=flags, tile=(tw, th), options=opts)
    n_active = sim.info().n_points
    print(f"[bench] uploaded and sorted; warmup {args.warmup} + one graph chunk", file=sys.stderr, flush=True)
    sim.run(0, args.warmup, graph=False)
    sim.run(args.warmup, 16, graph=True)     # graph capture and instantiation, outside the timer
    sim.sync()
    first = args.warmup + 16

    # timed region: graph replays of whole substeps
    sorts0 = sim.info().n_sorts
    t0 = time.perf_counter()
    sim.run(first, args.steps, graph=True)
    sim.sync()
    wall = time.perf_counter() - t0
    sorts_timed = sim.info().n_sorts - sorts0
    err, dis = _err_dis(sim)
    print(f"[bench] timed {args.steps} substeps in {wall:.3f}s; err {err:#x}", file=sys.stderr, flush=True)

    # per-kernel times: the next substeps, eagerly, with HIP events
    ev_first, ev_steps = first + args.steps, args.event_steps
    sim.reset_timings()
    sim.enable_timings(True)
    sim.sync()
    sim.run(ev_first, ev_steps, graph=False)
    sim.sync()
    tm = sim.timings()
    sim.enable_timings(False)

    nsub = max(tm[5], 1)
    ms_step = wall * 1000.0 / args.steps
    ms_p2g = tm[0] / nsub
    ms_g2p = tm[3] / nsub
    ms_sub_ev = tm[4] / nsub
    nodes = gx * gy
    # substeps of the event pass that store the complete point state (the last
    # of each 16-substep chunk) and that record P, Q
    chunks = [min(16, ev_steps - k) for k in range(0, ev_steps, 16)]
    every = scene.params.UpdateEveryNthStep
    n_rec = sum(1 for k in range(ev_first, ev_first + ev_steps) if every > 0 and k % every == 0)
    B = algorithmic_bytes(n_active, nodes, True, complete_frac=len(chunks) / ev_steps, rec_frac=n_rec / ev_steps)
    if pipelined:
        # one launch is one whole substep (the node update and G2P of substep n,
        # the P2G of n + 1 from registers): priced at the fused design's bytes,
        # 104 B per point + 33 B per node; the two-pass figure and the bytes this
        # design itself has to move are reported beside it
        dominant, ms_dom, kname = "fused", ms_g2p, "k_tile_pipeline"
    elif ms_g2p >= ms_p2g:
        dominant, ms_dom, kname = "g2p", ms_g2p, "k_g2p_tiles" if not args.naive else "k_g2p_naive"
    else:
        # the unfused P2G is the tile pass's P2G-only variant into the frames
        dominant, ms_dom, kname = "p2g", ms_p2g, ("k_p2g_naive" if args.naive else
                                                  "k_p2g_tiles_det" if args.det else "k_tile_pipeline<P2G>")
    achieved = B[dominant] / (ms_dom * 1e-3) / 1e9
    value = n_active * args.steps / wall
    # the dominant kernel's bytes per wall-clock substep (every kernel and gap
    # of the graph replays included) and per event-timed substep
    B_sub = B["fused"] if pipelined else B["substep"]
    wall_gbs = B_sub / (ms_step * 1e-3) / 1e9
    ev_gbs = B_sub / (ms_sub_ev * 1e-3) / 1e9
    own = two_pass_launch = None
    if pipelined:
        own_gbs = B["pipelined"] / (ms_dom * 1e-3) / 1e9
        own = {"bytes": B["pipelined"], "GBs": own_gbs, "frac": own_gbs / HBM_PEAK_GBS}
        tp = B["substep"] / (ms_dom * 1e-3) / 1e9
        two_pass_launch = {"bytes": B["substep"], "GBs": tp, "frac": tp / HBM_PEAK_GBS,
                           "note": "the two-pass substep's 168 B/point + 33 B/node priced on the same launch"}
    sim.close()

    secondary = {}
    if not args.no_secondary:
        secondary["batched"] = measure_batch(args.secondary_steps, autotune=False)
        secondary["surrogate"] = measure_surrogate(args.secondary_steps, args.surrogate_tol)

    cpu = None
    if args.cpu_seconds > 0:
        print("[bench] CPU baseline (FP64 oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scene, args.cpu_seconds, args.cpu_threads)

    out = {
        "metric": "particle-updates/s per substep at 32M particles (plain FP32 substep; surrogate leg beside)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"icefield_{gx}x{gy}: plain MPM, {gx}x{gy} nodes h=1 m, {n_active} points at 4/cell in a seeded "
                         "random polygonal ice field (~95%), GRF wind stress (mean 0.2 Pa, std 0.05 Pa, "
                         "corr 200 cells), lognormal cohesion (median 1e5 Pa, sigma 0.5), dt=1e-4 s"),
            "points": n_active, "nodes": nodes, "tile": args.tile,
            "path": "naive" if args.naive else ("sorted-deterministic" if args.det else
                                                ("sorted-unfused" if args.unfused else
                                                 ("sorted-pipelined-persistent" if args.persistent else
                                                  "sorted-pipelined"))),
            "sorts_in_timed_region": sorts_timed,
            "error_word": err,
        },
        "roofline": {
            "bound": "hbm", "kernel": kname,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(kname, f"{tw}x{th}", f"{gx}x{gy}"),
            "bytes_per_launch": B[dominant], "ms_per_launch": ms_dom,
            "substep_events": {"bytes": B_sub, "ms": ms_sub_ev, "GBs": ev_gbs, "frac": ev_gbs / HBM_PEAK_GBS},
            "substep_wall_clock": {"bytes": B_sub, "ms": ms_step, "GBs": wall_gbs, "frac": wall_gbs / HBM_PEAK_GBS},
            "p2g_ms": ms_p2g, "g2p_ms": ms_g2p, "event_substeps": ev_steps,
            "two_pass_bytes_same_launch": two_pass_launch,
            "own_design": own,
        },
        "cpu_baseline": cpu,
    }
    out.update(secondary)
    print(json.dumps(out), flush=True)
    return 0


def _timed_run(sim, steps: int, graph: bool = True) -> float:
    """One graph chunk of warm-up (capture, instantiation), then `steps`
    substeps of graph replays, timed (graph False: eager launches)."""
    sim.run(0, 16, graph=graph)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(16, steps, graph=graph)
    sim.sync()
    return time.perf_counter() - t0


def measure_batch(steps: int, autotune: bool = True, persistent: bool = False, resort_drift=None,
                  single_handle: bool = False) -> dict:
    """batch64_uneven: 64 independent domains (sides 128-1024 cells, 25-95 % ice)
    run (a) as one batched handle (one launch per phase, HIP-graph replays) and
    (b) one domain after another with the single-domain path, the reference's
    one-partition-at-a-time shape, each over the same `steps` substeps.
    single_handle skips (b), which creates one handle per domain (profiled
    runs use one handle)."""
    from icempm_2d_sp_amd import scenes
    from icempm_2d_sp_amd.gpu import BatchSimulation, Simulation
    t0 = time.time()
    doms = scenes.batch64_uneven()
    print(f"[bench] {len(doms)} domains, {sum(d.n for d in doms)} points, generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    steps = max(16, (steps // 16) * 16)
    opts = {"persistent": int(persistent)}
    if resort_drift is not None:
        opts["resort_drift"] = resort_drift
    b = BatchSimulation(doms, options=opts)
    tile, tune_ms = b.autotune(2) if autotune else ((16, 32), None)
    n_active = b.n_points()
    sorts0 = sum(b.info(d).n_sorts for d in range(len(doms)))
    wall_b = _timed_run(b, steps)
    sorts_b = sum(b.info(d).n_sorts for d in range(len(doms))) - sorts0
    b.close()
    print(f"[bench] batched: {steps} substeps in {wall_b:.3f}s (tile {tile})", file=sys.stderr, flush=True)
    wall_s = 0.0
    if not single_handle:
        for sc in doms:
            s1 = Simulation(sc)
            wall_s += _timed_run(s1, steps)
            s1.close()
        print(f"[bench] sequential: {wall_s:.3f}s", file=sys.stderr, flush=True)
    return {
        "workload": "batch64_uneven: 64 domains, sides log-uniform 128-1024 cells, 25-95% ice, 4 points/cell, "
                    "per-domain wind 0.05-0.3 Pa, plain FP32 substep",
        "points": n_active, "steps": steps, "tile": f"{tile[0]}x{tile[1]}", "sorts_in_timed_region": sorts_b,
        "value": n_active * steps / wall_b, "unit": "particle-updates/s", "ms_per_step": wall_b * 1000 / steps,
        "sequential_single_domain": None if single_handle else {"value": n_active * steps / wall_s,
                                                                 "ms_per_step": wall_s * 1000 / steps},
        "batched_over_sequential": None if single_handle else wall_s / wall_b,
    }


def bench_batch(args) -> int:
    """batch64_uneven (measure_batch), tile autotuned; value = the batched rate."""
    r = measure_batch(args.steps, autotune=True, persistent=args.persistent, resort_drift=args.resort_drift,
                      single_handle=args.single_handle)
    out = {
        "metric": "particle-updates/s per substep, 64 batched uneven domains (plain FP32 substep)",
        "value": r["value"], "unit": "particle-updates/s", "n_gpus": 1, "steps": r["steps"], "warmup": 16,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": r["workload"], "points": r["points"], "tile": r["tile"],
                   "sorts_in_timed_region": r["sorts_in_timed_region"]},
        "sequential_single_domain": r["sequential_single_domain"],
        "batched_over_sequential": r["batched_over_sequential"],
    }
    print(json.dumps(out), flush=True)
    return 0


def heldout_states(sim, first_step: int, steps: int = 4, n: int = 1 << 16):
    """Trial states of `steps` substeps after the training window (recorded,
    hash-sampled, at most n): states the network was not trained on."""
    sim.mlp_record(1, 0.0)
    sim.run(first_step, 1, graph=False)
    seen1, _ = sim.mlp_record_count()
    sim.mlp_record(n, min(1.0, 1.5 * n / max(seen1 * steps, 1)))
    sim.run(first_step + 1, steps, graph=False)
    x = sim.mlp_records(n)
    sim.mlp_record(0, 0.0)
    return x


def surrogate_accuracy(sim, x, tol: float) -> dict:
    """Stress error of the network against the FP64 exact projection on trial
    states x, and the acceptance at tol."""
    import numpy as np
    out = sim.mlp_eval(x, tol)
    ps, qs, pe, qe, viol, acc = (out[:, k].astype(np.float64) for k in range(6))
    err = np.hypot(ps - pe, qs - qe) / np.hypot(pe, qe)
    accepted = acc > 0.5
    err_dep = np.where(accepted, err, 0.0)   # fallback points take the exact map
    return {"network_median": float(np.median(err)), "network_p99": float(np.percentile(err, 99)),
            "network_max": float(err.max()), "violation_median": float(np.median(viol)),
            "accepted_fraction_heldout": float(accepted.mean()),
            "deployed_p99": float(np.percentile(err_dep, 99)), "deployed_max": float(err_dep.max()),
            "heldout_states": int(x.shape[0])}


def measure_surrogate(steps: int, tol: float, naive_leg: bool = False, single_handle: bool = False) -> dict:
    """batch64_compressed: 64 uneven domains in uniaxial compression at 1/s
    against the fixed wall at x_max under 1 Pa wind (most points project
    plastically every substep).  Times (a) plain batched substeps (exact
    return map), (b) batched substeps with the learned plastic projection and
    the exact fallback at `tol`, (c) the plain single-domain runs one after
    another, and with naive_leg (d) the reference's launch structure one
    domain after another.  The network is trained on the GPU first (not
    timed).  single_handle: (b) only, with eager launches, for profiling its
    kernels (one handle in the process and no HIP graphs)."""
    import numpy as np
    from icempm_2d_sp_amd import abi, scenes
    from icempm_2d_sp_amd.gpu import BatchSimulation, Simulation
    doms = scenes.batch64_compressed()
    steps = max(16, (steps // 16) * 16)
    print(f"[bench] {len(doms)} domains, {sum(d.n for d in doms)} points; {steps} substeps", file=sys.stderr,
          flush=True)
    pa = wall_plain = None
    if not single_handle:
        plain = BatchSimulation(doms)
        wall_plain = _timed_run(plain, steps)
        pa, _, _, _ = plain.download(0)
        plain.close()
        print(f"[bench] plain batched {wall_plain:.3f}s", file=sys.stderr, flush=True)

    surr = BatchSimulation(doms)
    n_active = surr.n_points()
    # the network is trained on the trial states of this workload's first 17
    # substeps (the 256x128 floe's run has none: it never yields), and judged
    # on those of the 4 substeps after them
    train = surr.train_on_recorded(0, 16, 10_000_000)
    x_test = heldout_states(surr, 17, 4)
    acc = surrogate_accuracy(surr, x_test, tol)
    print(f"[bench] trained on {train['pairs']} recorded pairs in {train['seconds']:.2f}s, loss "
          f"{train['loss_first']:.3g} -> {train['loss_last']:.3g}; held-out stress error median "
          f"{acc['network_median']:.2e} p99 {acc['network_p99']:.2e}", file=sys.stderr, flush=True)
    surr.reset_points()
    surr.mlp_deploy(True, tol)
    c0 = surr.mlp_counts()
    wall_surr = _timed_run(surr, steps, graph=not single_handle)
    c1 = surr.mlp_counts()
    n_acc, n_fb = c1[0] - c0[0], c1[1] - c0[1]   # warm-up chunk included: counts per substep below
    pb, _, err_surr, _ = surr.download(0)
    surr.close()
    print(f"[bench] surrogate batched {wall_surr:.3f}s, projections {n_acc}, fallbacks {n_fb}", file=sys.stderr,
          flush=True)
    dfe = wall_seq = None
    if not single_handle:
        ka = np.argsort(pa[abi.IDX_POINT].view(np.uint32), kind="stable")
        kb = np.argsort(pb[abi.IDX_POINT].view(np.uint32), kind="stable")
        dfe = float(np.abs(pa[abi.IDX_FE00:abi.IDX_FE00 + 4][:, ka].astype(np.float64) -
                           pb[abi.IDX_FE00:abi.IDX_FE00 + 4][:, kb]).max())
        wall_seq = 0.0
        for sc in doms:
            s1 = Simulation(sc)
            wall_seq += _timed_run(s1, steps)
            s1.close()
        print(f"[bench] plain sequential {wall_seq:.3f}s", file=sys.stderr, flush=True)

    def rate(wall):
        return None if wall is None else {"value": n_active * steps / wall, "ms_per_step": wall * 1000 / steps}
    out = {
        "workload": "batch64_compressed: 64 domains, sides log-uniform 128-1024 cells, 25-95% ice, uniaxial "
                    "compression at 1/s against the fixed wall at x_max, 1.0 Pa wind along +x, 4 points/cell",
        "points": n_active, "steps": steps, "surrogate_tol": tol, "dtype": "f32 state, bf16 MLP",
        "value": n_active * steps / wall_surr, "unit": "particle-updates/s", "ms_per_step": wall_surr * 1000 / steps,
        "plain_batched": rate(wall_plain),
        "plain_sequential": rate(wall_seq),
        "surrogate_batched_over_plain_sequential": None if wall_seq is None else wall_seq / wall_surr,
        "surrogate_batched_over_plain_batched": None if wall_plain is None else wall_plain / wall_surr,
        "ridge_capped_plastic_fraction_per_substep": (n_acc + n_fb) / (n_active * (steps + 16)),
        "surrogate_accepted_fraction": n_acc / max(n_acc + n_fb, 1),
        "training": train,
        "stress_error_vs_fp64": acc,
        "final_Fe_max_abs_diff_first_domain": dfe, "error_word": err_surr,
    }
    if naive_leg and not single_handle:
        # the reference's own launch structure (unsorted points, global atomics,
        # ICE_FLAG_NAIVE, eager launches), one domain after another
        naive_steps = 64
        wall_naive = 0.0
        for sc in doms:
            s1 = Simulation(sc, flags=sc.params.flags | abi.FLAG_NAIVE)
            s1.run(0, 16, graph=False)
            s1.sync()
            t0 = time.perf_counter()
            s1.run(16, naive_steps, graph=False)
            s1.sync()
            wall_naive += time.perf_counter() - t0
            s1.close()
        ms_naive = wall_naive * 1000 / naive_steps
        print(f"[bench] plain sequential, reference launch structure: {ms_naive:.3f} ms per substep",
              file=sys.stderr, flush=True)
        out["plain_sequential_reference_launch_structure"] = {
            "value": n_active * 1000 / ms_naive, "ms_per_step": ms_naive, "substeps": naive_steps,
            "note": "unsorted points, P2G with global atomics, separate update and G2P (ICE_FLAG_NAIVE)"}
        out["surrogate_batched_over_plain_sequential_reference_launch_structure"] = ms_naive / (wall_surr * 1000 / steps)
    return out


def bench_surrogate(args) -> int:
    """batch64_compressed (measure_surrogate, with the naive leg); value =
    the surrogate batched rate."""
    r = measure_surrogate(args.steps, args.surrogate_tol, naive_leg=True, single_handle=args.single_handle)
    out = {
        "metric": "particle-updates/s per substep, 64 batched uneven domains, surrogate + batched vs plain",
        "value": r["value"], "unit": "particle-updates/s", "n_gpus": 1, "steps": r["steps"], "warmup": 16,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": r["dtype"], "data": "synthetic",
        "config": {"workload": r["workload"], "points": r["points"], "surrogate_tol": r["surrogate_tol"]},
    }
    out.update({k: v for k, v in r.items() if k not in out and k not in ("workload", "points", "steps",
                                                                         "surrogate_tol", "dtype")})
    print(json.dumps(out), flush=True)
    return 0


def bench_compressed(args) -> int:
    """batch64_compressed, plain batched substep only: the fast-flow case of
    the sorted path (points cross cells every few substeps, domains re-sort
    between run chunks).  value = points * substeps / wall time over graph
    replays after one 16-substep warm-up chunk."""
    from icempm_2d_sp_amd import scenes
    from icempm_2d_sp_amd.gpu import BatchSimulation
    doms = scenes.batch64_compressed()
    steps = max(16, (args.steps // 16) * 16)
    opts = {"persistent": int(args.persistent)}
    if args.resort_drift is not None:
        opts["resort_drift"] = args.resort_drift
    b = BatchSimulation(doms, options=opts)
    n_active = b.n_points()
    b.run(0, 16, graph=True)
    b.sync()
    sorts0 = sum(b.info(d).n_sorts for d in range(len(doms)))
    t0 = time.perf_counter()
    b.run(16, steps, graph=True)
    b.sync()
    wall = time.perf_counter() - t0
    sorts = sum(b.info(d).n_sorts for d in range(len(doms))) - sorts0
    b.close()
    print(f"[bench] compressed batch: {steps} substeps in {wall:.3f}s, {sorts} domain sorts", file=sys.stderr,
          flush=True)
    out = {
        "metric": "particle-updates/s per substep, 64 batched uneven domains in uniaxial compression (plain FP32 substep)",
        "value": n_active * steps / wall, "unit": "particle-updates/s", "n_gpus": 1, "steps": steps, "warmup": 16,
        "ms_per_step": wall * 1000 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch64_compressed: the batch64_uneven domains in uniaxial compression at 1/s "
                               "against the fixed wall at x_max plus 1 Pa wind", "points": n_active, "tile": "16x32",
                   "domain_sorts_in_timed_region": sorts},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_frame(args) -> int:
    """The frame arrays (SnapshotManager::PrepareFrameArrays,
    snapshotmanager.cpp:623-765) of icefield_4096x2048 after 16 substeps:
    device time per frame (sort + splat + normalisation + region sums, small
    outputs), the whole call with every array copied to the host, and the
    oracle's serial host loop (the reference's structure) on the same state."""
    import numpy as np
    import oracle
    from icempm_2d_sp_amd import scenes
    from icempm_2d_sp_amd.gpu import Simulation
    scene = scenes.icefield_4096x2048()
    prm = scene.params
    gx, gy = prm.GridXTotal, prm.GridYTotal
    sim = Simulation(scene, tile=tuple(int(v) for v in args.tile.split("x")))
    n = sim.info().n_points
    colors = np.random.default_rng(1).integers(0, 1 << 24, n, dtype=np.uint32)
    sim.upload_point_colors(colors)
    sim.run(0, 15, graph=True)
    sim.substep(15 * prm.InitialTimeStep, True)
    sim.sync()
    sim.prepare_frame_arrays(16, want=("count", "forces_per_region"))   # warm-up (sorts the moved points)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.prepare_frame_arrays(16, want=("count", "forces_per_region"))
    t_dev = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    out = sim.prepare_frame_arrays(16)
    t_all = time.perf_counter() - t0
    pts, forces, err, _ = sim.download()
    sim.close()
    oracle.threads(1)
    t0 = time.perf_counter()
    ref = oracle.frame_arrays(pts, gx, gy, prm.ParticleMass, scene.status, colors, forces,
                              fscale=1.0 / (16 * prm.InitialTimeStep))
    t_cpu = time.perf_counter() - t0
    clear = np.abs(out["density"] - 0.1) >= 1e-4    # masks decided within rounding of the threshold aside
    dmax = max(float(np.abs(out[k][clear].astype(np.float64) - ref[k][clear]).max() /
                     max(np.abs(ref[k]).max(), 1e-30)) for k in ("density", "mass", "P", "Q"))
    outj = {
        "metric": "frame arrays (PrepareFrameArrays) per frame, 4096x2048 nodes, 31.8M particles",
        "value": 1000.0 * t_dev, "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
        "data": "synthetic", "config": {"workload": f"icefield_{gx}x{gy} after 16 substeps", "points": n},
        "device_ms_per_frame": 1000.0 * t_dev, "with_all_arrays_to_host_ms": 1000.0 * t_all,
        "cpu_baseline": {"value": 1000.0 * t_cpu, "unit": "ms", "cores": 1, "kind": "port",
                         "sample": "FP64 oracle restatement of the reference's serial host loop, whole state"},
        "max_rel_diff_vs_oracle": dmax, "error_word": err,
    }
    print(json.dumps(outj), flush=True)
    return 0


def _err_dis(sim):
    import ctypes as C
    from icempm_2d_sp_amd import abi
    err, dis = C.c_uint32(0), C.c_int32(0)
    abi.check(sim.lib.ice_download(sim.h, None, 0, abi.N_LINES, None, C.byref(err), C.byref(dis)), sim.h,
              "ice_download")
    return err.value, dis.value


if __name__ == "__main__":
    sys.exit(main())
