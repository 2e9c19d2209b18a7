#!/usr/bin/env python3
"""Benchmark of the cubff per-epoch hot path on one MI355X.

Workload: bff_noheads, 2^17 programs x 64 bytes, uniform SplitMix64 bytes from
seed 0, default mutation probability 1/4096 (P = 262144 / 2^30), at most 8192
steps per interaction, Fisher-Yates pairing every epoch.  One "step" is one
epoch: device pairing + mutate + evaluate + write back over the whole soup.

Prints one JSON line.  `value` is interpreted instructions per second (the
reference's ops: executed non-no-op instructions) over the timed epochs, with
the soup resident in HBM.  The roofline object prices the epoch kernel
(k_epoch, one launch per epoch) with the frozen integer work model

    W_epoch = (N/2) * 128 * C_mut + sum(steps) * C_step,
    C_mut = 32 lane-ops per mutation draw, C_step = 8 lane-ops per BFF step,

where steps counts every interpreter loop iteration including no-ops, over
the kernel's average launch time from HIP events on the context stream,
against 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 int32 lane-ops/s.  `traffic`
is the measured HBM bytes per launch (committed rocprofv3 PMC summary) and
`hbm` the algorithmic bytes per launch (per pair slot: 128 gathered, 128
scattered, 8 of pairing, 8 of per-slot results) over the launch time,
against 8 TB/s, as evidence.  The kernel is bound by neither: an epoch lasts
as long as its longest interaction.

The CPU baseline runs the reference build from oracle/_ref (compiled from the
reference sources) on a bounded sample of the same workload on the host's
cores; without it, the oracle's OpenMP port is timed instead.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("interpreted instr/s + epochs/s, 2^17×64B bff_noheads; "
          "batched+sched vs plain")
PEAK_LANE_OPS = 256 * 64 * 2.4e9
PEAK_HBM = 8e12          # bytes/s, MI355X HBM3E
# Epoch kernel per pair slot: 2 x 64 B gathered + 2 x 64 B scattered +
# s[2k..2k+1] + ops, steps (+ the two programs' owner words when deferring).
BYTES_PER_SLOT = {"loop": 272, "defer": 280, "dataflow": 280}
C_MUT, C_STEP = 32, {"bff": 8, "forth": 10, "subleq": 10, "rsubleq4": 12}


def c_step(lang: str) -> int:
    if lang.startswith("bff"):
        return C_STEP["bff"]
    if lang.startswith("forth"):
        return C_STEP["forth"]
    return C_STEP[lang]


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(lang: str, n: int, seed: int, prob: int, first: int, count: int,
                 budget_s: float):
    """Times the reference CPU path (or the oracle port) on the host over the
    same epochs as the GPU value, first..first+count-1.

    The reference's RunSimulation reports, at every callback, its rate over
    the epochs since the previous callback with the callback's own work
    (soup copy, brotli, byte tables) excluded (common_language.h:496-512,
    :583-585).  A callback interval that divides both first-1 and count
    makes those windows tile the GPU's epochs exactly; the run stops after
    the last of them, and is repeated from epoch 0 while the wall-clock
    budget lasts (or cut short inside the window if one run overruns it)."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    last = first + count - 1
    from oracle import refimpl
    if refimpl.available():
        interval = math.gcd(first - 1, count) if first >= 1 else max(count - 1, 1)
        ops = 0
        secs = 0.0
        reps = 0
        covered = last
        t_start = time.perf_counter()
        while True:
            prev = {"ops": 0, "epoch": 0}
            got = {"ops": 0, "secs": 0.0, "end": first - 1}
            t_run = time.perf_counter()

            def cb(st):
                # This callback closes the window of epochs prev.epoch..st.epoch-1.
                lo, hi = prev["epoch"], int(st.epoch) - 1
                d_ops = int(st.total_ops) - prev["ops"]
                prev["ops"], prev["epoch"] = int(st.total_ops), int(st.epoch)
                if lo >= first and hi <= last:
                    got["ops"] += d_ops
                    got["secs"] += d_ops / (float(st.mops_s) * 1e6)
                    got["end"] = hi
                    if hi >= last:
                        return True
                    # One run longer than the whole budget is cut short.
                    return reps == 0 and time.perf_counter() - t_run > budget_s
                return False

            refimpl.run_simulation(lang, n, seed, cb, mutation_prob=prob,
                                   callback_interval=interval)
            ops += got["ops"]
            secs += got["secs"]
            reps += 1
            covered = min(covered, got["end"])
            if got["end"] < last or time.perf_counter() - t_start + (
                    time.perf_counter() - t_run) > budget_s:
                break
        return {"value": ops / secs, "unit": "instr/s", "cores": threads, "kind": "reference",
                "epochs_per_s": reps * (covered - first + 1) / secs,
                "cpu": cpu_model(),
                "sample": (f"reference RunSimulation (oracle/_ref, g++ -O3 -fopenmp, "
                           f"{threads} OpenMP threads) on the same workload and the same "
                           f"epochs {first}..{covered} from seed {seed}, run {reps} time(s); "
                           f"time from the reference's own per-callback rate (callbacks "
                           f"every {interval} epochs, their work excluded)")}
    from oracle import oracle as O
    S = O.Soup(lang, n, seed, prob, nthreads=threads)
    for _ in range(first):
        S.step()
    t0 = time.perf_counter()
    ops = 0
    e = 0
    while e < count and time.perf_counter() - t0 < budget_s:
        ops += S.step()[0]
        e += 1
    dt = time.perf_counter() - t0
    return {"value": ops / dt, "unit": "instr/s", "cores": threads, "kind": "port",
            "epochs_per_s": e / dt, "cpu": cpu_model(),
            "sample": f"oracle OpenMP port, epochs {first}..{first + e - 1} of the same workload"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_k_epoch_defer_lean_bff_noheads_2p17.md")


def traffic_per_launch():
    """HBM bytes per epoch-kernel launch from the committed PMC summary (FETCH_SIZE
    + WRITE_SIZE per dispatch, each scaled by the same counter's reading for a
    plain copy of known size), or None when it has not been measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


BATCH_SIZES = [4096, 16384, 65536, 131072]


def batched_config():
    """64 bff_noheads soups: seeds 0..63, sizes cycling 2^12, 2^14, 2^16,
    2^17, mutation probability log-spaced 1e-5 .. 1e-3."""
    S = 64
    ns = [BATCH_SIZES[k % 4] for k in range(S)]
    probs = [int(round(10 ** (-5 + 2 * k / 63) * (1 << 30))) for k in range(S)]
    return ["bff_noheads"] * S, ns, list(range(S)), probs


AUTOTUNE_ROUNDS = 3


def batched_start(warmup):
    """First timed epoch of the batched legs: after the warm-up and the
    longest launch-shape sweep (rounds x 12 shapes x two-epoch runs), so
    that every leg times the same epochs."""
    return warmup + AUTOTUNE_ROUNDS * 12 * 2


def run_batched(capi, sched, epochs, warmup, autotune, spec=None):
    langs, ns, seeds, probs = spec if spec is not None else batched_config()
    with capi.Batch(langs, ns) as b:
        b.init(seeds, probs)
        tuned = None
        # Warm-up first (the learned schedule trains on its first epoch), then
        # the launch-shape sweep on live epochs, then untimed epochs up to the
        # common first timed epoch, then the timed epochs.
        b.run_epochs(0, warmup, sched)
        b.sync()
        e0 = warmup
        if autotune:
            wg, tpl, table = b.autotune(e0, AUTOTUNE_ROUNDS, sched)
            tuned = {"wg": wg, "tapes_per_lane": tpl, "ms_per_epoch": table}
            e0 = b.next_epoch()
        start = batched_start(warmup)
        if e0 < start:
            b.run_epochs(e0, start - e0, sched)
            e0 = start
        b.sync()
        b.reset_counters()
        tot, it = b.run_epochs_timed(e0, epochs, sched)
        ops = steps = 0
        for i in range(len(ns)):
            o, _, st = b.read_counters(i)
            ops += o
            steps += st
        wg, tpl = b.get_launch()
        # The next epochs replayed from the captured epoch graph (interpreter
        # rounds; a replayed epoch does not defer).
        b.reset_counters()
        gtot, _ = b.run_epochs_timed(e0 + epochs, epochs, sched, use_graph=True)
        gops = sum(b.read_counters(i)[0] for i in range(len(ns)))
        graph = {"first_epoch": e0 + epochs, "epochs_timed": epochs,
                 "ms_per_epoch": gtot / epochs, "instr_per_s": gops / (gtot / 1e3)}
        out = {"instr_per_s": ops / (tot / 1e3), "ms_per_epoch": tot / epochs,
               "first_epoch": e0,
               "interp_ms_per_epoch": it / epochs, "epochs_timed": epochs,
               "total_programs": sum(ns), "soups": len(ns), "wg": wg,
               "tapes_per_lane": tpl, "steps": steps, "ops": ops, "autotune": tuned,
               "graph_replay": graph}
        if sched == capi.SCHED_LEARNED:
            tr, loss, buckets = b.sched_stats()
            out["sched"] = {"trainings": tr, "last_loss": loss,
                            "bucket_slots": [int(x) for x in buckets]}
        return out


def run_single_same_soups(capi, spec, first, epochs):
    """The soups of a batch run one at a time through the plain single-soup
    path (the library's default multi-epoch path per soup), same languages,
    sizes, seeds, mutation probabilities and epochs first..first+epochs-1 as
    the batched leg's timed span.  Times add up over the soups."""
    langs, ns, seeds, probs = spec
    tot = 0.0
    ops = 0
    for lang, n, seed, prob in zip(langs, ns, seeds, probs):
        with capi.Soup(lang, n) as g:
            g.init_programs(capi.epoch_seed(seed, 0))
            if first:
                g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, first, prob)
            g.sync()
            g.reset_counters()
            t, _ = g.run_epochs_timed(capi.PAIR_SHUFFLE, seed, first, epochs, prob)
            o, _, _ = g.read_counters(with_steps=True)
        tot += t
        ops += o
    return {"instr_per_s": ops / (tot / 1e3), "ms_per_epoch": tot / epochs,
            "first_epoch": first, "epochs_timed": epochs, "ops": ops}


def same_soups_leg(capi, spec, learned, plain):
    """learned batched vs the plain single-soup kernel on the same soups."""
    single = run_single_same_soups(capi, spec, learned["first_epoch"], learned["epochs_timed"])
    return {"plain_single_same_soups": single,
            "same_work": single["ops"] == learned["ops"],
            "learned_batched_vs_plain_single_same_soups":
                learned["instr_per_s"] / single["instr_per_s"],
            "plain_batched_vs_plain_single_same_soups":
                plain["instr_per_s"] / single["instr_per_s"]}


def run_soup_timed(capi, lang, n, seed, prob, warmup, epochs, mode=None, pattern=None,
                   flow=None):
    """Single-soup epochs timed with HIP events; returns rates."""
    mode = capi.PAIR_SHUFFLE if mode is None else mode
    with capi.Soup(lang, n) as g:
        if flow is not None:
            g.set_pipeline(flow)
        g.init_programs(capi.epoch_seed(seed, 0))
        if pattern is not None:
            g.set_interactions(*pattern)
        g.run_epochs(mode, seed, 0, warmup, prob)
        g.sync()
        g.reset_counters()
        if mode == capi.PAIR_2D:
            # The 2-D matching loops on the host until every edge is decided,
            # so it is timed by the wall clock around synchronised epochs.
            t0 = time.perf_counter()
            g.run_epochs(mode, seed, warmup, epochs, prob)
            g.sync()
            tot, kern = (time.perf_counter() - t0) * 1e3, float("nan")
        else:
            tot, kern = g.run_epochs_timed(mode, seed, warmup, epochs, prob)
        ops, runs, steps = g.read_counters(with_steps=True)
    return {"instr_per_s": ops / (tot / 1e3), "epochs_per_s": epochs / (tot / 1e3),
            "ms_per_epoch": tot / epochs,
            "kernel_ms_per_epoch": None if kern != kern else kern / epochs,
            "epochs_timed": epochs}


BROTLI_LIB = "/usr/lib/x86_64-linux-gnu/libbrotlienc.so.1"


def brotli_size(data) -> int:
    """BrotliEncoderCompress(2, 24, BROTLI_MODE_GENERIC) of the soup, the
    reference's entropy measurement (common_language.h:502-507), through the
    system libbrotlienc (ctypes releases the GIL for the call)."""
    import ctypes
    import numpy as np
    lib = ctypes.CDLL(BROTLI_LIB)
    lib.BrotliEncoderMaxCompressedSize.restype = ctypes.c_size_t
    lib.BrotliEncoderMaxCompressedSize.argtypes = [ctypes.c_size_t]
    lib.BrotliEncoderCompress.restype = ctypes.c_int
    lib.BrotliEncoderCompress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                          ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t),
                                          ctypes.c_void_p]
    src = np.ascontiguousarray(data, dtype=np.uint8)
    cap = lib.BrotliEncoderMaxCompressedSize(src.size)
    dst = np.empty(cap, dtype=np.uint8)
    size = ctypes.c_size_t(cap)
    if not lib.BrotliEncoderCompress(2, 24, 0, src.size, src.ctypes.data, ctypes.byref(size),
                                     dst.ctypes.data):
        raise RuntimeError("BrotliEncoderCompress failed")
    return int(size.value)


def run_large_soup(capi, epochs):
    """The large-soup run: bff_noheads, 2^22 programs, seed 7, P = 257698
    (p = 0.00024); every epoch is followed by the device self-rep check of
    all programs, and every 64 epochs the soup is copied back and brotli-
    compressed on a host thread while the next epochs run."""
    from concurrent.futures import ThreadPoolExecutor
    n, seed, P4 = 1 << 22, 7, 257698
    first = 2
    with capi.Soup("bff_noheads", n) as g, ThreadPoolExecutor(1) as pool:
        g.init_programs(capi.epoch_seed(seed, 0))
        g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, first, P4)
        g.selfrep(capi.epoch_seed(seed, first - 1), n)  # first-use allocation, untimed
        g.sync()
        g.reset_counters()
        ep_s = sr_s = 0.0
        best = []
        pending = []
        t0 = time.perf_counter()
        for e in range(first, first + epochs):
            t1 = time.perf_counter()
            g.run_epochs(capi.PAIR_SHUFFLE, seed, e, 1, P4)
            g.sync()
            t2 = time.perf_counter()
            scores = g.selfrep(capi.epoch_seed(seed, e), n)
            t3 = time.perf_counter()
            ep_s += t2 - t1
            sr_s += t3 - t2
            best.append(int(scores.max()))
            if e % 64 == 0:
                pending.append((e, pool.submit(brotli_size, g.read_soup())))
        brotli = [(e, f.result()) for e, f in pending]
        total = time.perf_counter() - t0
        ops, _, _ = g.read_counters(with_steps=True)
    return {
        "workload": (f"bff_noheads, 2^22 programs, seed 7, P=257698 (p=0.00024), epochs "
                     f"{first}..{first + epochs - 1}, self-rep check of all programs after every "
                     f"epoch, brotli of the soup every 64 epochs on a host thread"),
        "epochs_timed": epochs, "instr_per_s": ops / ep_s, "ms_per_epoch": ep_s * 1e3 / epochs,
        "selfrep_ms_per_epoch": sr_s * 1e3 / epochs,
        "epochs_per_s_with_selfrep": epochs / total, "total_s": total,
        "brotli_bpb": {str(e): b * 8 / (n * 64) for e, b in brotli},
        "max_selfrep_score": max(best) if best else None}


PROBE = os.path.join(ROOT, "cubff_amd", "lib", "interp_probe")


def epoch_tapes(capi, lang, n, seed, prob, epoch):
    """The 128-byte tapes of epoch `epoch`'s pairs (before that epoch's
    mutation), the soup run there on the GPU."""
    import numpy as np
    with capi.Soup(lang, n) as g:
        g.init_programs(capi.epoch_seed(seed, 0))
        g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, epoch, prob)
        soup = g.read_soup().reshape(n, 64)
        g.run_epochs(capi.PAIR_SHUFFLE, seed, epoch, 1, prob)
        s, _ = g.read_shuffle()
    return np.ascontiguousarray(np.concatenate([soup[s[0::2]], soup[s[1::2]]], axis=1))


def interp_probe(capi, lang, n, seed, prob, epoch, budget=768):
    """Per-command cost of the BFF interpreters on the pairs of a real epoch:
    their tapes go to cubff_amd/lib/interp_probe, which times each
    interpreter per wavefront with the shader clock."""
    import subprocess
    import tempfile
    if not os.path.exists(PROBE):
        return None
    tapes = epoch_tapes(capi, lang, n, seed, prob, epoch)
    with tempfile.NamedTemporaryFile(suffix=".tapes") as f:
        f.write(tapes.tobytes())
        f.flush()
        out = subprocess.run([PROBE, f.name, str(budget)], capture_output=True, text=True,
                             timeout=120, check=True).stdout
    r = json.loads(out)
    r["workload"] = f"{lang}, {n} programs, seed {seed}: the {n // 2} pairs of epoch {epoch}"
    return r


def run_other_configs(capi, epochs, large_epochs):
    """The other benchmark configurations, each on one GPU."""
    from cubff_amd.patterns import grid_pattern
    out = {}
    P = 1 << 18
    if large_epochs > 0:
        out["large_soup"] = run_large_soup(capi, large_epochs)
    if epochs <= 0:
        return out
    langs = {}
    for lang in ("bff", "bff8", "forth", "forthtrivial", "subleq", "rsubleq4"):
        # One launch per epoch, the multi-epoch dataflow kernel, and the
        # default (which picks the dataflow kernel for the Forth family).
        loop = run_soup_timed(capi, lang, 131072, 1, P, 4, epochs, flow=capi.FLOW_OFF)
        flow = run_soup_timed(capi, lang, 131072, 1, P, 4, epochs, flow=capi.FLOW_ON)
        auto = run_soup_timed(capi, lang, 131072, 1, P, 4, epochs)
        auto["loop_ms_per_epoch"] = loop["ms_per_epoch"]
        auto["dataflow_ms_per_epoch"] = flow["ms_per_epoch"]
        langs[lang] = auto
    grid = grid_pattern(512, 256, 2)
    g_loop = run_soup_timed(capi, "bff_noheads", 512 * 256, 1, P, 4, epochs, mode=capi.PAIR_2D,
                            pattern=grid, flow=capi.FLOW_OFF)
    g_auto = run_soup_timed(capi, "bff_noheads", 512 * 256, 1, P, 4, epochs, mode=capi.PAIR_2D,
                            pattern=grid)
    g_auto["loop_ms_per_epoch"] = g_loop["ms_per_epoch"]
    langs["bff_noheads_grid_512x256"] = g_auto
    out["all_interpreters"] = {
        "workload": "2^17 programs, seed 1, Fisher-Yates pairing; plus bff_noheads on a "
                    "512x256 grid with radius-2 neighbour pairing", "langs": langs}
    # Learned scheduler on mixed languages.
    langs5 = ["bff" if s % 2 == 0 else "forth" for s in range(100, 132)]
    ns5 = [65536] * 32
    seeds5 = list(range(100, 132))
    probs5 = [P] * 32
    res = {}
    for name, sched in (("plain", capi.SCHED_PLAIN), ("learned_sched", capi.SCHED_LEARNED)):
        res[name] = run_batched(capi, sched, epochs, 8, autotune=True,
                                spec=(langs5, ns5, seeds5, probs5))
    res["learned_vs_plain_batched"] = (res["learned_sched"]["instr_per_s"] /
                                       res["plain"]["instr_per_s"])
    res.update(same_soups_leg(capi, (langs5, ns5, seeds5, probs5), res["learned_sched"],
                              res["plain"]))
    res["workload"] = ("32 soups x 2^16 programs, seeds 100..131, bff on even seeds and "
                       "forth on odd seeds, Fisher-Yates")
    out["mixed_sched"] = res
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256, help="timed epochs")
    ap.add_argument("--warmup", type=int, default=16, help="untimed epochs")
    ap.add_argument("--lang", default="bff_noheads")
    ap.add_argument("--num", type=int, default=131072)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall-clock budget of the CPU baseline sample (0 = skip)")
    ap.add_argument("--batched-epochs", type=int, default=32,
                    help="timed epochs of the 64-soup batched runs (0 = skip)")
    ap.add_argument("--config-epochs", type=int, default=16,
                    help="timed epochs of the other configurations (0 = skip)")
    ap.add_argument("--large-epochs", type=int, default=256,
                    help="timed epochs of the 2^22-program run with the self-rep check every "
                         "epoch (0 = skip)")
    ap.add_argument("--pipeline", choices=["auto", "loop", "dataflow", "defer"], default="auto",
                    help="multi-epoch path of the single-soup run (auto: the library's choice "
                         "for the language)")
    ap.add_argument("--interp-probe", type=int, default=1,
                    help="time the BFF interpreters per command on a real epoch's pairs "
                         "(cubff_amd/lib/interp_probe; 0 = skip)")
    ap.add_argument("--write-epoch-tapes", default="",
                    help="only write the tapes of epoch 60's pairs of the single-soup workload "
                         "to this file (input of cubff_amd/lib/interp_probe)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB "
                         "(four launches), the reference for HBM byte counters")
    args = ap.parse_args()
    if args.calibrate_copy_mib > 0:
        import torch
        nbytes = args.calibrate_copy_mib << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": 4}))
        return
    if args.write_epoch_tapes:
        from cubff_amd import capi
        with open(args.write_epoch_tapes, "wb") as f:
            f.write(epoch_tapes(capi, args.lang, args.num, args.seed, 1 << 18, 60).tobytes())
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the epoch runs on a single MI355X and "
              "this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from cubff_amd import capi
    prob = 1 << 18  # round(1/4096 * 2^30), main.cc:183, :226
    n, seed, lang = args.num, args.seed, args.lang
    modes = {"auto": capi.FLOW_AUTO, "loop": capi.FLOW_OFF, "dataflow": capi.FLOW_ON,
             "defer": capi.FLOW_DEFER}
    with capi.Soup(lang, n) as g:
        g.set_pipeline(modes[args.pipeline])
        g.init_programs(capi.epoch_seed(seed, 0))
        g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, args.warmup, prob)
        g.sync()
        g.reset_counters()
        _, _, d0 = g.flow_stats()
        t0 = time.perf_counter()
        total_ms, kern_ms = g.run_epochs_timed(capi.PAIR_SHUFFLE, seed, args.warmup,
                                               args.steps, prob)
        g.sync()
        wall = time.perf_counter() - t0
        ops, runs, steps = g.read_counters(with_steps=True)
        items, claims, d1 = g.flow_stats()
    deferred = d1 - d0
    path = ("defer" if deferred > 0 or args.pipeline == "defer" else
            "dataflow" if items > 0 else "loop")
    loop = None
    if path != "loop":
        # The same epochs with one launch per epoch, for comparison.
        loop = run_soup_timed(capi, lang, n, seed, prob, args.warmup, args.steps,
                              flow=capi.FLOW_OFF)

    K = args.steps
    dt = total_ms / 1e3
    work = (n // 2) * 128 * C_MUT * K + steps * c_step(lang)
    # On the deferring path the work is shared by the epoch kernel and the
    # resolver running beside it on a second stream, so the rate is taken
    # over the whole epoch span (both kernels and the end-of-run drain), not
    # over the epoch kernel's launches alone.
    basis_ms = total_ms if path == "defer" else kern_ms
    lane_ops = work / (basis_ms / 1e3)
    algo_bytes = (n // 2) * BYTES_PER_SLOT[path]
    achieved = algo_bytes / (basis_ms / 1e3 / K)
    # The BFF languages with the standard jump logic run the word-table
    # interpreter's epoch kernel; the others the byte-tape one.
    defer_kernel = ("k_epoch_defer_lean" if lang in ("bff", "bff_noheads", "bff_noheads_4bit",
                                                     "bff_perm") else "k_epoch_defer")
    line = {
        "metric": METRIC,
        "value": ops / dt,
        "unit": "instr/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": total_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": (f"{lang}, {n} programs x 64 B, seed {seed}, mutation "
                         f"P={prob}/2^30, 8192 max steps, Fisher-Yates pairing, "
                         f"epochs {args.warmup}..{args.warmup + K - 1}"),
            "lang": lang, "num_programs": n, "seed": seed,
            "parallelism": "single soup, one lane per pair slot",
            "path": {"loop": "one launch per epoch (k_epoch)",
                     "defer": f"{defer_kernel} per epoch (command budget 512) + resolver "
                              "kernel on a second stream",
                     "dataflow": "multi-epoch dataflow kernel (k_flow)"}[path],
        },
        "epochs_per_s": K / dt,
        "interp_steps_per_s": steps / dt,
        "ops_per_pair": ops / max(runs, 1),
        "wall_s": wall,
        "roofline": {
            "bound": "valu",
            "achieved": lane_ops / 1e12,
            "peak": PEAK_LANE_OPS / 1e12,
            "unit": "T int32 lane-ops/s",
            "frac": lane_ops / PEAK_LANE_OPS,
            # Measured for the default workload only (bff_noheads, 2^17 programs,
            # deferring epoch kernel).
            "traffic": (traffic_per_launch() if (lang, n, path) == ("bff_noheads", 131072, "defer")
                        else None),
            "kernel": {"loop": "k_epoch (gather + mutate + evaluate + scatter)",
                       "defer": f"{defer_kernel} (gather + mutate + evaluate up to the "
                                "budget + scatter)",
                       "dataflow": "k_flow"}[path],
            "kernel_ms_avg": kern_ms / K,
            "time_basis": ("epoch span (epoch kernel + concurrent resolver + drain), "
                           "ms_per_step" if path == "defer" else "kernel_ms_avg"),
            "work_per_launch": work / K,
            "model": "W = (N/2)*128*32 + sum(steps)*C_step",
            "limiter": "latency of the longest interaction of the epoch",
            "hbm": {"achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": achieved / PEAK_HBM,
                    "algorithmic_bytes_per_launch": algo_bytes},
        },
        "cpu_baseline": None,
        "deferred_per_epoch": deferred / K,
        "loop_one_launch_per_epoch": loop,
    }
    if args.interp_probe and lang.startswith("bff"):
        line["interp_probe"] = interp_probe(capi, lang, n, seed, prob, 60)
    if args.batched_epochs > 0:
        plain = run_batched(capi, capi.SCHED_PLAIN, args.batched_epochs, 8, autotune=True)
        learned = run_batched(capi, capi.SCHED_LEARNED, args.batched_epochs, 8, autotune=True)
        line["batched"] = {
            "workload": ("64 bff_noheads soups, seeds 0..63, sizes cycling 2^12/2^14/2^16/"
                         "2^17 programs, mutation log-spaced 1e-5..1e-3, Fisher-Yates"),
            "plain": plain, "learned_sched": learned,
            "learned_vs_plain_batched": learned["instr_per_s"] / plain["instr_per_s"],
        }
        line["batched"].update(same_soups_leg(capi, batched_config(), learned, plain))
    if args.config_epochs > 0 or args.large_epochs > 0:
        line["configs"] = run_other_configs(capi, args.config_epochs, args.large_epochs)
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(lang, n, seed, prob, args.warmup, args.steps,
                                            args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
