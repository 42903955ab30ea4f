"""Benchmark: sensor MP/s for Opposed + AMaZE + RL on one MI355X.

Workload: the full-sensor frame, 6048x4024 (24.3 MP) RGGB u16 14-bit (black
512, white 16383), seeded synthetic scene with 2 % clipped rectangles and 5 %
saturated Gaussian highlight blobs; Opposed highlight reconstruction, AMaZE
demosaic, masked Richardson-Lucy capture sharpening (sigma 0.7, amount 0.5,
noise protection 0.5, 10 iterations); f32 camera RGB out.

One step = one develop of that frame from its u16 mosaic resident in HBM to
the three resident f32 RGB planes (stats readback and stream sync included).
The timed region brackets exactly `--steps` steps with a synchronize on
both sides. The JSON line also carries the roofline of the dominant kernel
(timed live with HIP events on the library's stream) and the CPU baseline:
the oracle's C++ restatement of the same chain on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-served", action="store_true")
    p.add_argument("--batch", action="store_true",
                   help="run the batched workload instead: 32 mosaics of 12-61 MP with mixed CFA "
                        "phases, demosaic algorithms and RL iterations in one graph launch")
    p.add_argument("--batch-frames", type=int, default=32)
    p.add_argument("--serve", action="store_true",
                   help="run the serving workload instead: Poisson arrivals over 64 resident "
                        "24-45 MP frames, against the plain serial path on the same schedule")
    p.add_argument("--serve-requests", type=int, default=400)
    p.add_argument("--serve-rate", type=float, default=40.0, help="mean arrivals per second")
    p.add_argument("--c5", action="store_true",
                   help="102 MP 16-bit frames mixed 1:3 with 24 MP ones: stage + optics + "
                        "area reduction to a 2560 px long edge")
    p.add_argument("--c5-frames", type=int, default=4,
                   help="with --c5: use the first N frames of the 1:3 mix (1 = the 102 MP frame)")
    p.add_argument("--detail", action="store_true",
                   help="the 24 MP develop with the whole detail stage: false colour (2 steps), "
                        "chroma and luminance noise reduction (0.5) and RL")
    p.add_argument("--decode", action="store_true",
                   help="RAW ingest: the 24 MP mosaic as Sony compressed ARW, decoded on the GPU")
    p.add_argument("--batch-lanes", type=int, default=0, help="0: autotune")
    p.add_argument("--served-lanes", type=int, default=4)
    p.add_argument("--served-requests", type=int, default=60)
    p.add_argument("--served-poisson", type=int, default=200,
                   help="requests of the Poisson serving leg in the default line (0: skip)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run only a plain device copy of 1 GiB (read) into 1 GiB (write), the "
                        "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    p.add_argument("--rocprof-summary", metavar="RESULTS_DB",
                   help="print the per-kernel table of a rocprofv3 --kernel-trace --stats "
                        "results database and exit (no GPU needed)")
    p.add_argument("--traffic-summary", nargs=4,
                   metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                   help="print per-kernel HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE "
                        "and WRITE_SIZE counter CSVs of this bench and of --calibrate-copy, "
                        "scaled by the copy kernel's known bytes, and exit (no GPU needed)")
    p.add_argument("--counter-summary", metavar="COUNTER_CSV",
                   help="print per-kernel means of a rocprofv3 --pmc VALUBusy LdsUtil TA_BUSY_avr "
                        "MemUnitStalled SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_WAIT_INST_ANY "
                        "SQ_WAVE_CYCLES counter CSV of this bench and exit (no GPU needed)")
    return p.parse_args()


# Per-kernel HBM traffic measured with rocprofv3 counters (--traffic-summary
# output); the live roofline reports the dominant kernel's row when present.
TRAFFIC_FILE = ROOT / "profiles" / "traffic_develop_24mp.txt"
COUNTERS_FILE = ROOT / "profiles" / "counters_develop_24mp.txt"
COPY_BYTES = 1 << 30


def short_kernel_name(name):
    """`rhd::(anonymous namespace)::cap_iterate<3, false>(...)` -> `cap_iterate<3,false>`."""
    import re
    m = re.search(r"::(\w+)(<[^<>()]*>)?\(", name)
    return (m.group(1) + (m.group(2) or "").replace(" ", "")) if m else name.split("(")[0][:48]


def rocprof_summary(db):
    import sqlite3
    con = sqlite3.connect(db)
    rows = list(con.execute(
        "select name, total_calls, total_duration, average, percentage from top_kernels"))
    print(f"{'kernel':28s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s} {'%':>6s}")
    for name, calls, total, avg, pct in rows:
        # top_kernels holds microseconds.
        print(f"{short_kernel_name(name):28s} {calls:6d} {total / 1e3:10.3f} {avg / 1e3:9.4f} "
              f"{pct:6.2f}")


def counter_means(path, counter):
    """{kernel: (dispatches, mean counter value)} from a rocprofv3 counter CSV."""
    import csv
    acc = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            k = short_kernel_name(row["Kernel_Name"])
            n, total = acc.get(k, (0, 0.0))
            acc[k] = (n + 1, total + float(row["Counter_Value"]))
    return {k: (n, total / n) for k, (n, total) in acc.items()}


def traffic_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv):
    fetch = counter_means(fetch_csv, "FETCH_SIZE")
    write = counter_means(write_csv, "WRITE_SIZE")
    copy_f = counter_means(copy_fetch_csv, "FETCH_SIZE")["__amd_rocclr_copyBuffer"][1]
    copy_w = counter_means(copy_write_csv, "WRITE_SIZE")["__amd_rocclr_copyBuffer"][1]
    sf, sw = COPY_BYTES / (copy_f * 1024), COPY_BYTES / (copy_w * 1024)
    print(f"# copy kernel: FETCH_SIZE {copy_f:.1f} KiB, WRITE_SIZE {copy_w:.1f} KiB per dispatch "
          f"for {COPY_BYTES} bytes each way")
    print(f"# scale: fetch x{sf:.4f}, write x{sw:.4f} (bytes per counted byte)")
    print(f"{'kernel':28s} {'dispatches':>10s} {'fetch_bytes':>14s} {'write_bytes':>14s} "
          f"{'hbm_bytes':>14s}")
    for k, (n, f) in fetch.items():
        fb = f * 1024 * sf
        wb = write.get(k, (0, 0.0))[1] * 1024 * sw
        print(f"{k:28s} {n:10d} {fb:14.0f} {wb:14.0f} {fb + wb:14.0f}")


def counter_summary(path):
    cols = ["VALUBusy", "LdsUtil", "TA_BUSY_avr", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD"]
    means = {c: counter_means(path, c) for c in cols + ["SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"]}
    kernels = sorted(means["SQ_WAVE_CYCLES"], key=lambda k: -means["SQ_WAVE_CYCLES"][k][1])
    print(f"{'kernel':26s} {'VALUBusy':>9s} {'LdsUtil':>8s} {'TA_busy':>10s} {'VALU_insts':>12s} "
          f"{'VMEM_rd':>10s} {'wait/wave':>9s}")
    for k in kernels:
        v = [means[c].get(k, (0, float("nan")))[1] for c in cols]
        wait = means["SQ_WAIT_INST_ANY"].get(k, (0, 0.0))[1]
        cycles = means["SQ_WAVE_CYCLES"][k][1]
        print(f"{k:26s} {v[0]:9.1f} {v[1]:8.1f} {v[2]:10.3g} {v[3]:12.3g} {v[4]:10.3g} "
              f"{(wait / cycles if cycles else 0.0):9.2f}")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed traffic summary."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        parts = line.split()
        if len(parts) == 5 and not line.startswith("#") and parts[0] == kernel:
            return float(parts[4])
    return None


def measured_valu_busy(kernel):
    """VALUBusy (% of the kernel's time the vector ALUs issue) of `kernel`
    from the committed counter summary."""
    if not COUNTERS_FILE.exists():
        return None
    for line in COUNTERS_FILE.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 2 and not line.startswith("#") and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def calibrate_copy(nbytes=1 << 30):
    """A plain f32 copy of known size: its FETCH_SIZE and WRITE_SIZE per
    dispatch, against nbytes, give the counters' scale on this device."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "dispatches": 3}))


def workload():
    from rohditor_amd import sensor, synth
    spec = synth.full_sensor_24mp()
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=10)
    options = sensor.RenderOptions(demosaic=sensor.DemosaicAlgorithm.AMAZE)
    return spec, frame, recipe, options


def _cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with its family / model /
    stepping (a virtualised host may report only a generic name)."""
    fields = {}
    try:
        for line in open("/proc/cpuinfo"):
            key, _, value = line.partition(":")
            key = key.strip()
            if key in ("model name", "cpu family", "model", "stepping") and key not in fields:
                fields[key] = value.strip()
            if key == "flags" and "flags" not in fields:
                flags = value.split()
                fields["flags"] = " ".join(f for f in ("avx2", "fma", "avx512f") if f in flags)
    except OSError:
        return "unknown"
    return (f"{fields.get('model name', 'unknown')} (family {fields.get('cpu family', '?')}, "
            f"model {fields.get('model', '?')}, stepping {fields.get('stepping', '?')}; "
            f"{fields.get('flags', '')})")


def _cpu_chain(frame, recipe, options, threads):
    """Seconds for the oracle's normalize + Opposed + AMaZE + RL on `frame`."""
    from oracle import oracle
    from rohditor_amd import sensor
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
    c = desc.crop
    t0 = time.perf_counter()
    norm = oracle.normalize(frame.mosaic, desc.sensor_pattern, desc.black, *desc.black_repeat,
                            desc.white, (c.x, c.y, c.width, c.height))
    _, norm, _, _, _ = oracle.opposed(norm, desc.crop_pattern, desc.detection_levels,
                                      desc.neutral_ceiling)
    _, rgb = oracle.demosaic(norm, desc.crop_pattern, desc.demosaic_balance, "amaze",
                             threads=threads)
    ceil = sensor.capture_ceilings(recipe, desc.demosaic_balance)
    oracle.capture(rgb, recipe.capture.amount, recipe.capture.radius,
                   recipe.capture.noise_protection, ceil, iterations=recipe.capture_iterations,
                   threads=threads)
    return time.perf_counter() - t0, c.width * c.height / 1e6


def cpu_baseline(frame, recipe, options):
    """The oracle's restatement of the same chain, built -O3 -march=native
    for this host (contraction off, so its results are the checker's): the
    full frame on every host core (1 warm-up, median of 5), and a 1536x1024
    crop of it on one core (median of 3)."""
    from oracle import oracle
    from rohditor_amd import synth
    threads = oracle.cpu_threads()
    build = "-O3 -march=native -ffp-contract=off"
    try:
        native = oracle.native_build()
    except Exception:  # no compiler on this host: the -O2 checker build
        native, build = None, "-O2 -ffp-contract=off (checker build)"
    crop = synth.frame_from_mosaic(np.ascontiguousarray(frame.mosaic[:1024, :1536]),
                                   frame.info.cfa, float(frame.info.black_levels.values[0]),
                                   float(frame.info.white_levels[0]))
    with (oracle.using(native) if native else _nullcontext()):
        _cpu_chain(frame, recipe, options, threads)
        runs = [_cpu_chain(frame, recipe, options, threads) for _ in range(5)]
        single = [_cpu_chain(crop, recipe, options, 1) for _ in range(3)]
    dt = float(np.median([r[0] for r in runs]))
    mp = runs[0][1]
    dt1 = float(np.median([r[0] for r in single]))
    return {"value": round(mp / dt, 3), "unit": "MP/s", "cores": threads, "kind": "port",
            "build": build,
            "sample": f"one full {frame.info.width}x{frame.info.height} frame, "
                      f"normalize+Opposed+AMaZE+RL({recipe.capture_iterations}), median of 5 "
                      f"after 1 warm-up: {dt:.2f} s",
            "single_core": {"value": round(single[0][1] / dt1, 3), "unit": "MP/s", "cores": 1,
                            "sample": f"{crop.info.width}x{crop.info.height} crop of the frame, "
                                      f"median of 3: {dt1:.2f} s"},
            "cpu_model": _cpu_model()}


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not supported "
              "(frames shard by frame id across GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.rocprof_summary:
        rocprof_summary(args.rocprof_summary)
        return
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary)
        return
    if args.counter_summary:
        counter_summary(args.counter_summary)
        return
    if args.batch:
        batch_bench(args)
        return
    if args.serve:
        serve_bench(args)
        return
    if args.c5:
        c5_bench(args)
        return
    if args.detail:
        detail_bench(args)
        return
    if args.decode:
        decode_bench(args)
        return
    from rohditor_amd import _lib, sensor

    spec, frame, recipe, options = workload()
    proc = sensor.GpuSensorProcessor(0)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    resident = proc.upload(frame)
    mp = desc.crop.width * desc.crop.height / 1e6

    def step(metrics=None):
        src = proc.develop_resident(resident, desc, detail, metrics=metrics)
        src.release()

    for _ in range(args.warmup):
        step()
    # Per-kernel table: the same steps with HIP events around every launch.
    # Those events cost about 0.3 ms per develop, so they stay out of the
    # timed region below.
    proc.lib.rhd_profile_reset(proc.ctx)
    proc.lib.rhd_profile_enable(proc.ctx, 1)
    for _ in range(args.steps):
        step()
    proc.lib.rhd_synchronize(proc.ctx)
    proc.lib.rhd_profile_enable(proc.ctx, 0)
    prof = profile_table(proc)
    # Dominant kernel: the one with the most GPU time among those with a
    # host-known algorithmic byte count (the Opposed ring-entry kernels'
    # sizes are known only on the device).
    top = max(prof, key=lambda k: prof[k][0])
    dominant = max((k for k in prof if prof[k][2] > 0), key=lambda k: prof[k][0])

    # Timed region: in its last `sampled` steps the dominant kernel's launches
    # carry HIP events, on the stream they are launched on (each event pair
    # adds about 10 us to the develop, so the other steps carry none).
    sampled = min(2, args.steps)
    proc.lib.rhd_profile_reset(proc.ctx)
    proc.lib.rhd_profile_only(proc.ctx, dominant.encode())
    proc.lib.rhd_synchronize(proc.ctx)
    lat = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i == args.steps - sampled:
            proc.lib.rhd_profile_enable(proc.ctx, 1)
        s0 = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s0)
    proc.lib.rhd_synchronize(proc.ctx)
    total = time.perf_counter() - t0
    proc.lib.rhd_profile_enable(proc.ctx, 0)
    proc.lib.rhd_profile_only(proc.ctx, None)

    ms_per_step = total * 1e3 / args.steps
    value = mp * args.steps / total
    dom_ms, dom_launches, dom_bytes = profile_table(proc)[dominant]
    avg_ms = dom_ms / max(dom_launches, 1)
    achieved = dom_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # rocprof names the dominant kernel's instantiation (cap_iterate<3,false>
    # for sigma 0.7: radius 3, estimate stored); the traffic file is keyed by
    # that name.
    traffic = None
    valu_busy = None
    for name in (dominant, dominant + "<3>", dominant + "<3,false>"):
        traffic = traffic if traffic is not None else measured_traffic(name)
        valu_busy = valu_busy if valu_busy is not None else measured_valu_busy(name)
    # The roofline priced here is HBM's; the kernel's own limiter comes from
    # its counters: vector-ALU issue when VALUBusy exceeds the HBM fraction.
    frac = achieved / HBM_PEAK_GBPS
    limiter = "valu" if valu_busy is not None and valu_busy / 100.0 > frac else "hbm"
    roofline = {"bound": "hbm", "limiter": limiter,
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": None if traffic is None else int(traffic),
                "traffic_source": None if traffic is None else str(TRAFFIC_FILE.relative_to(ROOT)),
                "kernel": dominant, "avg_launch_ms": round(avg_ms, 4),
                "launches_per_step": dom_launches / sampled,
                "sampled_steps": sampled,
                "algorithmic_bytes_per_launch": int(dom_bytes),
                "top_kernel": top, "top_kernel_ms_per_step": round(prof[top][0] / args.steps, 4),
                # The kernel's actual limiter: vector-ALU issue, from a --pmc pass.
                "valu_busy_pct": valu_busy,
                "valu_busy_source": None if valu_busy is None else str(COUNTERS_FILE.relative_to(ROOT))}
    stage = {"stage_GBps": round(mp * 1e6 * 14 / (ms_per_step * 1e-3) / 1e9, 1),
             "stage_frac": round(mp * 1e6 * 14 / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}

    served_line = None if args.no_served else served(frame, desc, detail, value, args)
    poisson = None
    if not args.no_served and args.served_poisson > 0:
        sp, pp = poisson_legs(args, args.served_poisson, args.serve_rate)
        poisson = {"served": sp, "plain_serial": pp, "requests": args.served_poisson,
                   "rate_per_s": args.serve_rate, "lanes": args.served_lanes,
                   "workload": poisson_workload(args.serve_rate)}
    cpu = None if args.no_cpu_baseline else cpu_baseline(frame, recipe, options)
    lat_ms = np.array(lat) * 1e3
    line = {
        "metric": "sensor MP/s (AMaZE+Opposed+RL) + p50/p99 develop latency, served vs plain",
        "value": round(value, 2), "unit": "MP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "full-sensor 6048x4024 RGGB 14-bit, Opposed+AMaZE+RL(sigma 0.7, "
                               "10 iterations), plain single-frame develop from a resident mosaic",
                   "width": spec.width, "height": spec.height, "megapixels": round(mp, 3),
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                       "p99": round(float(np.percentile(lat_ms, 99)), 3)},
        "roofline": roofline,
        "stage_roofline": stage,
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in sorted(
            prof.items(), key=lambda kv: -kv[1][0])},
        "served": served_line,
        "served_poisson": poisson,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def served(frame, desc, detail, plain_value, args):
    """The same develop through the serving loop (include/rhd.h rhd_server_*):
    a closed loop keeping 2 x lanes requests in flight (its latency includes
    the queueing behind them), each on its own
    resident frame id (so newest-wins never coalesces them), no base cache
    (every request runs the whole stage). Latency is submit -> RGB ready."""
    from rohditor_amd import sensor
    lanes, frames = args.served_lanes, 4 * args.served_lanes
    in_flight, total = 2 * lanes, args.served_requests
    proc = sensor.GpuSensorProcessor(0)
    server = sensor.GpuServer(proc, lanes=lanes)
    for f in range(frames):
        server.put_frame(f, frame)
    mp = desc.crop.width * desc.crop.height / 1e6

    def run(n):
        pending, nxt, done = [], 0, 0
        while done < n:
            while len(pending) < in_flight and nxt < n:
                pending.append(server.submit(nxt % frames, desc, detail))
                nxt += 1
            r = server.wait(pending.pop(0), keep=False)
            if not r.ok:
                raise RuntimeError(f"served develop failed with status {r.status}")
            done += 1

    run(2 * in_flight)  # warm-up: every lane's scratch and the device pool
    server.reset_stats()
    t0 = time.perf_counter()
    run(total)
    wall = time.perf_counter() - t0
    st = server.stats()
    server.close()
    proc.close()
    value = mp * total / wall
    return {"value": round(value, 2), "unit": "MP/s", "vs_plain": round(value / plain_value, 3),
            "latency_ms": {"p50": round(st.p50_ms, 3), "p99": round(st.p99_ms, 3),
                           "mean": round(st.mean_ms, 3)},
            "requests": total, "lanes": lanes, "in_flight": in_flight,
            "workload": "closed loop over resident copies of the same frame and recipe, "
                        "no base cache"}


BATCH_SIZES = [(4240, 2832), (6048, 4024), (7008, 4672), (8256, 5504), (9504, 6336)]


def batch_bench(args):
    """The batched workload: `--batch-frames` mosaics with sizes drawn from
    12, 24, 33, 45 and 61 MP, CFA phase from RGGB/BGGR/GRBG, demosaic from
    AMaZE/RCD/MHC and RL iterations from 0..20 (seeded), Opposed highlight
    recovery, developed by one graph launch; against the same frames
    developed one after another on the plain path."""
    from rohditor_amd import sensor, synth
    rng = np.random.default_rng(0x52480300)
    n = args.batch_frames
    sizes = rng.integers(0, len(BATCH_SIZES), n)
    cfas = rng.integers(0, 3, n)
    algos = rng.integers(0, 3, n)
    iters = rng.integers(0, 21, n)
    # One synthetic scene per size (seed 0x52480300 + size index); frames of
    # a size share its samples and differ in CFA phase and recipe.
    scenes = {}
    for k in sorted(set(sizes.tolist())):
        w, h = BATCH_SIZES[k]
        scenes[k] = synth.mosaic(synth.SceneSpec(width=w, height=h, seed=0x52480300 + k,
                                                 blob_fraction=0.05))
    proc = sensor.GpuSensorProcessor(0, budget_bytes=200 << 30)
    residents, descs, details, mp = [], [], [], 0.0
    algo_of = [sensor.DemosaicAlgorithm.AMAZE, sensor.DemosaicAlgorithm.RCD,
               sensor.DemosaicAlgorithm.MHC]
    for i in range(n):
        w, h = BATCH_SIZES[sizes[i]]
        cfa = ["RGGB", "BGGR", "GRBG"][cfas[i]]
        frame = synth.frame_from_mosaic(scenes[sizes[i]], cfa)
        recipe = sensor.EditRecipe(
            highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
            capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                             noise_protection=0.5),
            capture_iterations=int(iters[i]))
        desc = sensor.SensorDevelopmentDescription.from_frame(
            frame, recipe, sensor.RenderOptions(demosaic=algo_of[algos[i]]))
        residents.append(proc.upload(frame))
        descs.append(desc)
        details.append(sensor.detail_desc(recipe, desc.demosaic_balance))
        mp += w * h / 1e6
    # Plain path: one develop after another.
    for i in range(min(2, n)):  # warm-up
        proc.develop_resident(residents[i], descs[i], details[i]).release()
    t0 = time.perf_counter()
    for i in range(n):
        proc.develop_resident(residents[i], descs[i], details[i]).release()
    plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = sensor.GpuBatch(proc, residents, descs, details, lanes=args.batch_lanes)
    build_s = time.perf_counter() - t0
    batch.run()  # warm-up
    runs = []
    for _ in range(max(args.steps // 10, 3)):
        t0 = time.perf_counter()
        statuses, gpu_ms = batch.run()
        runs.append(time.perf_counter() - t0)
        if any(s != 0 for s in statuses):
            raise RuntimeError(f"batched develop failed: {statuses}")
    wall = float(np.median(runs))
    line = {
        "metric": "batched sensor MP/s (Opposed + AMaZE/RCD/MHC + RL 0-20), one graph launch",
        "value": round(mp / wall, 2), "unit": "MP/s", "n_gpus": 1, "higher_is_better": True,
        "frames": n, "megapixels": round(mp, 2), "ms_per_batch": round(wall * 1e3, 3),
        "plain_serial_value": round(mp / plain_s, 2), "vs_plain": round(plain_s / wall, 3),
        "lanes": batch.lanes, "autotuned": args.batch_lanes == 0,
        "build_and_tune_s": round(build_s, 3), "runs": len(runs), "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "32 seeded mosaics, sizes from {4240x2832, 6048x4024, 7008x4672, "
                               "8256x5504, 9504x6336}, CFA RGGB/BGGR/GRBG, demosaic "
                               "AMaZE/RCD/MHC, Opposed, RL sigma 0.7 with 0-20 iterations"},
    }
    batch.close()
    for r in residents:
        r.release()
    proc.close()
    print(json.dumps(line))


SERVE_SIZES = [(6048, 4024), (7008, 4672), (8256, 5504)]


def poisson_legs(args, n, rate):
    """The serving workload: 64 resident frames of 24, 33 and 45 MP; `n`
    Poisson arrivals (`rate` per second, seeded); 70 % re-develops of a frame
    already developed that change its highlight method or demosaic
    algorithm, 30 % first develops of a frame; RL (sigma 0.7, 10 iterations)
    throughout. The server runs with newest-wins, admission by device bytes
    and the demosaiced-base cache. The same schedule is then replayed through
    the plain path, one develop at a time in arrival order."""
    import threading
    from rohditor_amd import sensor, synth
    rng = np.random.default_rng(0x52480444)
    frames_n = 64
    scenes = [synth.mosaic(synth.SceneSpec(width=w, height=h, seed=0x52480400 + k,
                                           blob_fraction=0.05))
              for k, (w, h) in enumerate(SERVE_SIZES)]
    cfas = ["RGGB", "BGGR", "GRBG"]
    frames = [synth.frame_from_mosaic(scenes[i % 3], cfas[(i // 3) % 3]) for i in range(frames_n)]
    methods = [sensor.HighlightMethod.OPPOSED, sensor.HighlightMethod.CLIP]
    algos = [sensor.DemosaicAlgorithm.AMAZE, sensor.DemosaicAlgorithm.RCD,
             sensor.DemosaicAlgorithm.MHC]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    developed, current, reqs = [], {}, []
    for i in range(n):
        if developed and rng.random() < 0.7:
            f = int(developed[rng.integers(0, len(developed))])
            m, a = current[f]
            if rng.random() < 0.5:
                m = 1 - m
            else:
                a = (a + 1 + int(rng.integers(0, 2))) % 3
        else:
            f = int(rng.integers(0, frames_n))
            m, a = 0, 0
            if f not in current:
                developed.append(f)
        current[f] = (m, a)
        recipe = sensor.EditRecipe(
            highlights=sensor.HighlightAdjustments(method=methods[m]),
            capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                             noise_protection=0.5),
            capture_iterations=10)
        desc = sensor.SensorDevelopmentDescription.from_frame(
            frames[f], recipe, sensor.RenderOptions(demosaic=algos[a]))
        reqs.append((float(arrivals[i]), f, desc, sensor.detail_desc(recipe, desc.demosaic_balance)))
    mp_of = lambda d: d.crop.width * d.crop.height / 1e6

    # Served.
    proc = sensor.GpuSensorProcessor(0, budget_bytes=200 << 30)
    server = sensor.GpuServer(proc, lanes=args.served_lanes, admission_bytes=48 << 30,
                              base_cache_bytes=16 << 30)
    for f in range(frames_n):
        server.put_frame(f, frames[f])
    for _, f, desc, detail in reqs[:4]:  # warm-up: lane scratch and the device pool
        server.wait(server.submit(f, desc, detail), keep=False)
    server.reset_stats()
    ids = [None] * n
    t_start = time.perf_counter()

    def submitter():
        for i, (at, f, desc, detail) in enumerate(reqs):
            delay = t_start + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids[i] = server.submit(f, desc, detail)

    th = threading.Thread(target=submitter)
    th.start()
    done_mp, statuses = 0.0, {}
    for i in range(n):
        while ids[i] is None:
            time.sleep(0.0005)
        r = server.wait(ids[i], keep=False)
        statuses[r.status] = statuses.get(r.status, 0) + 1
        if r.ok:
            done_mp += r.megapixels
    th.join()
    st = server.stats()
    served_line = {"value": round(st.megapixels / (st.span_ms * 1e-3), 2),
                   "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
                   "mean_ms": round(st.mean_ms, 3), "completed": int(st.completed),
                   "coalesced": int(st.coalesced), "cancelled": int(st.cancelled),
                   "base_hits": int(st.base_hits), "failed": int(st.failed),
                   "span_s": round(st.span_ms * 1e-3, 3)}
    server.close()

    # Plain: the same schedule, one develop at a time in arrival order.
    residents = {}
    lat, busy = [], 0.0
    for f in sorted(set(r[1] for r in reqs)):
        residents[f] = proc.upload(frames[f])
    for _, f, desc, detail in reqs[:4]:  # warm-up
        proc.develop_resident(residents[f], desc, detail).release()
    t_start = time.perf_counter()
    total_mp = 0.0
    for at, f, desc, detail in reqs:
        now = time.perf_counter() - t_start
        if now < at:
            time.sleep(at - now)
        t0 = time.perf_counter()
        proc.develop_resident(residents[f], desc, detail).release()
        t1 = time.perf_counter()
        busy += t1 - t0
        lat.append((t1 - t_start - at) * 1e3)
        total_mp += mp_of(desc)
    span = time.perf_counter() - t_start
    for r in residents.values():
        r.release()
    proc.close()
    lat = np.array(lat)
    plain_line = {"value": round(total_mp / span, 2), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                  "p99_ms": round(float(np.percentile(lat, 99)), 3),
                  "busy_mp_per_s": round(total_mp / busy, 2), "completed": n,
                  "span_s": round(span, 3)}
    return served_line, plain_line


def serve_bench(args):
    served_line, plain_line = poisson_legs(args, args.serve_requests, args.serve_rate)
    n = args.serve_requests
    line = {
        "metric": "served sensor MP/s and p50/p99 develop latency vs plain serial GPU path",
        "value": served_line["value"], "unit": "MP/s", "n_gpus": 1, "higher_is_better": True,
        "requests": n, "rate_per_s": args.serve_rate, "lanes": args.served_lanes,
        "served": served_line, "plain_serial": plain_line,
        "latency_p99_vs_plain": round(plain_line["p99_ms"] / max(served_line["p99_ms"], 1e-9), 3),
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": poisson_workload(args.serve_rate)},
    }
    print(json.dumps(line))


def poisson_workload(rate):
    return (f"Poisson arrivals ({rate}/s, seed 0x52480444) over 64 resident frames of "
            "6048x4024, 7008x4672 and 8256x5504; 70 % re-develops changing highlight "
            "(Opposed/Clip) or demosaic (AMaZE/RCD/MHC), 30 % first develops; RL sigma 0.7, "
            "10 iterations; newest-wins, 48 GiB admission, 16 GiB base cache")


C5_SIZES = [(11648, 8736), (6048, 4024), (6048, 4024), (6048, 4024)]


def c5_plan(w, h, rng):
    """Poly5 (k1, k2 in +-0.05), quartic PA vignetting, linear TCA 1 +- 5e-4."""
    from rohditor_amd import optics as opt
    return opt.resolve_execution(
        w, h, camera_crop=1.0, focal=24.0,
        distortion=opt.Poly5(*rng.uniform(-0.05, 0.05, 2)),
        vignetting=opt.Pa(float(rng.uniform(-0.4, -0.1)), float(rng.uniform(0.0, 0.1)), 0.0),
        tca=opt.TcaLinear(*(1.0 + rng.uniform(-5e-4, 5e-4, 2))))


def c5_bench(args):
    """One 11648x8736 16-bit frame (black 1024, white 65535, seed 0x52480005)
    and three 24 MP ones (seeds 0x52480006..8), each developed (Opposed +
    AMaZE + RL(10)), lens-corrected and area-reduced to a 2560 px long edge.
    Timed twice: the whole chain per frame, and with the developed bases kept
    on the device, the spatial step alone (an optics or preview-size edit)."""
    from rohditor_amd import optics as opt
    from rohditor_amd import sensor, synth
    rng = np.random.default_rng(0x52480005)
    proc = sensor.GpuSensorProcessor(0, budget_bytes=200 << 30)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=10)
    options = sensor.RenderOptions(demosaic=sensor.DemosaicAlgorithm.AMAZE)
    frames = []
    for k, (w, h) in enumerate(C5_SIZES[:max(1, args.c5_frames)]):
        spec = synth.SceneSpec(w, h, 0x52480005 + k, black=1024.0, white=65535.0,
                               blob_fraction=0.05)
        frame = synth.frame(spec)
        desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
        detail = sensor.detail_desc(recipe, desc.demosaic_balance)
        plan = c5_plan(w, h, rng)
        target = opt.preview_dimensions(w, h, 2560)
        frames.append((proc.upload(frame), desc, detail, plan, target, w * h / 1e6))
        print(f"c5: frame {k} {w}x{h} ready, preview {target}, scale {plan.scale}",
              file=sys.stderr, flush=True)
    mp = sum(f[5] for f in frames)

    def chain(f):
        resident, desc, detail, plan, target, _ = f
        src = proc.develop_resident(resident, desc, detail)
        small, ms = proc.reduce_preview(src, plan, target)
        src.release()
        small.release()
        return ms

    for f in frames:  # warm-up
        chain(f)
    steps = max(args.steps // 4, 3)
    t0 = time.perf_counter()
    for _ in range(steps):
        for f in frames:
            chain(f)
    chain_s = (time.perf_counter() - t0) / steps

    # Cached bases: the spatial step alone, profiled per kernel.
    bases = [proc.develop_resident(f[0], f[1], f[2]) for f in frames]
    spatial_gpu = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        for k, (f, base) in enumerate(zip(frames, bases)):
            small, ms = proc.reduce_preview(base, f[3], f[4])
            small.release()
            spatial_gpu.setdefault(k, []).append(ms)
    spatial_s = (time.perf_counter() - t0) / steps
    # Per-kernel times of the 102 MP frame's spatial step.
    proc.lib.rhd_profile_reset(proc.ctx)
    proc.lib.rhd_profile_enable(proc.ctx, 1)
    for _ in range(steps):
        small, _ = proc.reduce_preview(bases[0], frames[0][3], frames[0][4])
        small.release()
    proc.lib.rhd_profile_enable(proc.ctx, 0)
    prof = profile_table(proc)
    for b in bases:
        b.release()

    big_ms = float(np.median(spatial_gpu[0]))
    rest = [ms for k in range(1, len(frames)) for ms in spatial_gpu[k]]
    small_ms = float(np.median(rest)) if rest else None
    # Algorithmic bytes of the spatial step: the corrected RGB read once
    # (12 B/px) and the reduced RGB written once.
    w, h = C5_SIZES[0]
    tw, th = frames[0][4]
    alg = w * h * 12 + tw * th * 12
    dom = "opt_reduce_h"
    dom_ms, dom_launches, dom_bytes = prof.get(dom, (0.0, 1, 0))
    per_launch_ms = dom_ms / max(dom_launches, 1)
    line = {
        "metric": "C5 sensor MP/s: stage + optics + area reduction, 102 MP and 24 MP mixed 1:3",
        "value": round(mp / chain_s, 2), "unit": "MP/s", "n_gpus": 1, "higher_is_better": True,
        "steps": steps, "frames_per_step": len(frames), "megapixels_per_step": round(mp, 3),
        "ms_per_step": round(chain_s * 1e3, 3), "dtype": "f32", "data": "synthetic",
        "spatial_only": {
            "mp_per_s": round(mp / spatial_s, 1), "ms_per_step": round(spatial_s * 1e3, 3),
            "gpu_ms_102mp": round(big_ms, 3),
            "gpu_ms_24mp": round(small_ms, 3) if small_ms is not None else None,
            "GBps_102mp": round(alg / (big_ms * 1e-3) / 1e9, 1),
            "frac_102mp": round(alg / (big_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
        "roofline": {"bound": "hbm", "kernel": dom, "avg_launch_ms": round(per_launch_ms, 4),
                     "algorithmic_bytes_per_launch": int(dom_bytes),
                     "achieved": round(dom_bytes / (per_launch_ms * 1e-3) / 1e9, 1)
                     if per_launch_ms > 0 else None,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s"},
        "kernels_ms_102mp": {k: round(v[0] / steps, 4) for k, v in sorted(
            prof.items(), key=lambda kv: -kv[1][0])},
        "config": {"workload": "11648x8736 16-bit (black 1024, white 65535, seed 0x52480005) + "
                               "3 x 6048x4024 16-bit (seeds 0x52480006-8), 5% highlight blobs; "
                               "Opposed + AMaZE + RL(sigma 0.7, 10 iterations); Poly5 + quartic "
                               "PA vignetting + linear TCA; exact area reduction to 2560 px",
                   "parallelism": "single GPU"},
    }
    for f in frames:
        f[0].release()
    proc.close()
    print(json.dumps(line))


NR_KERNELS = ("nr_est_split", "nr_b3_h", "nr_est_band", "nr_est_tiles", "nr_keys", "nr_sort",
              "nr_gather", "nr_select", "nr_select_hist", "nr_fit", "nr_chroma_level",
              "nr_chroma_last", "nr_luma_h", "nr_luma_v", "nr_luma_energy_h", "nr_luma_shrink",
              "nr_luma_last")


def detail_bench(args):
    """The 24 MP frame of the headline workload developed with the whole
    detail stage (crates/core/src/pipeline/detail.rs:210-307): noise profiles
    of the demosaiced frame, false colour (2 median steps), chroma and
    luminance noise reduction (strength 0.5, detail 0.5), RL(10). Timed next
    to the headline develop without them; per-kernel times from a profiled
    pass. The CPU baseline runs the oracle's detail stage on a 1536x1024 crop
    of the same demosaiced frame."""
    from rohditor_amd import sensor
    spec, frame, base_recipe, options = workload()
    recipe = sensor.EditRecipe(
        highlights=base_recipe.highlights, capture=base_recipe.capture,
        capture_iterations=base_recipe.capture_iterations, false_color_steps=2,
        noise_reduction=sensor.NoiseReduction(chroma=0.5, luminance=0.5, luminance_detail=0.5))
    proc = sensor.GpuSensorProcessor(0)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    plain = sensor.detail_desc(base_recipe, desc.demosaic_balance)
    resident = proc.upload(frame)
    mp = desc.crop.width * desc.crop.height / 1e6

    def timed(dd):
        for _ in range(args.warmup):
            proc.develop_resident(resident, desc, dd).release()
        proc.lib.rhd_synchronize(proc.ctx)
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s0 = time.perf_counter()
            proc.develop_resident(resident, desc, dd).release()
            lat.append(time.perf_counter() - s0)
        proc.lib.rhd_synchronize(proc.ctx)
        return (time.perf_counter() - t0) / args.steps, np.array(lat) * 1e3

    full_s, lat = timed(detail)
    plain_s, _ = timed(plain)
    print(f"detail: full {full_s * 1e3:.3f} ms, plain {plain_s * 1e3:.3f} ms", file=sys.stderr,
          flush=True)
    proc.lib.rhd_profile_reset(proc.ctx)
    proc.lib.rhd_profile_enable(proc.ctx, 1)
    for _ in range(args.steps):
        proc.develop_resident(resident, desc, detail).release()
    proc.lib.rhd_profile_enable(proc.ctx, 0)
    prof = profile_table(proc)
    nr = {k: v for k, v in prof.items() if k in NR_KERNELS}
    nr_ms = sum(v[0] for v in nr.values()) / args.steps
    dom = max((k for k in nr if nr[k][2] > 0), key=lambda k: nr[k][0])
    dom_ms, dom_launches, dom_bytes = nr[dom]
    avg_ms = dom_ms / max(dom_launches, 1)
    achieved = dom_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        base = proc.develop_resident(resident, desc, plain)
        rgb = base.read_rgb()[:1024, :1536].copy()
        base.release()
        cpu = detail_cpu_baseline(rgb, detail)
    resident.release()
    proc.close()
    line = {
        "metric": "sensor MP/s with the whole detail stage (false colour + chroma/luma NR + RL)",
        "value": round(mp / full_s, 2), "unit": "MP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(full_s * 1e3, 3), "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "plain_ms_per_step": round(plain_s * 1e3, 3),
        "noise_reduction_gpu_ms_per_step": round(nr_ms, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "roofline": {"bound": "hbm", "kernel": dom, "avg_launch_ms": round(avg_ms, 4),
                     "launches_per_step": dom_launches / args.steps,
                     "sampled_steps": args.steps,
                     "algorithmic_bytes_per_launch": int(dom_bytes),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4)},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in sorted(
            prof.items(), key=lambda kv: -kv[1][0])},
        "cpu_baseline": cpu,
        "config": {"workload": "full-sensor 6048x4024 RGGB 14-bit, Opposed+AMaZE, false colour "
                               "2 steps, chroma NR 0.5, luminance NR 0.5 (detail 0.5), "
                               "RL(sigma 0.7, 10 iterations)", "parallelism": "single GPU"},
    }
    print(json.dumps(line))


def decode_bench(args):
    """RAW ingest of the 24 MP frame: its mosaic encoded as Sony compressed
    ARW (8 bits per pixel, tone curve SONY_CURVE), then per step the whole
    ingest as the engine would run it: container walk on the host, strip
    upload (1 B/px instead of the 2 B/px of a decoded u16 mosaic) and the
    decode kernel. Timed next to uploading the decoded u16 mosaic. The CPU
    baseline is the oracle's single-threaded decode of the same strip."""
    from rohditor_amd import sensor, synth
    spec = synth.full_sensor_24mp()
    mosaic = synth.mosaic(spec)
    strip = synth.encode_arw2(synth.sony_codes(mosaic))
    data = synth.arw_file(strip, spec.width, spec.height)
    proc = sensor.GpuSensorProcessor(0)
    mp = spec.width * spec.height / 1e6

    def ingest():
        layout = sensor.arw_probe(data)
        m, ms = proc.decode_arw(data, layout)
        m.release()
        return ms

    decoded = None
    for _ in range(args.warmup):
        ingest()
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gpu_ms.append(ingest())
    ingest_s = (time.perf_counter() - t0) / args.steps
    m, _ = proc.decode_arw(data)
    decoded = m.read(spec.width, spec.height)
    m.release()
    frame = synth.frame_from_mosaic(decoded, "RGGB")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proc.upload(frame).release()
    upload_s = (time.perf_counter() - t0) / args.steps
    proc.lib.rhd_profile_reset(proc.ctx)
    proc.lib.rhd_profile_enable(proc.ctx, 1)
    for _ in range(args.steps):
        ingest()
    proc.lib.rhd_profile_enable(proc.ctx, 0)
    prof = profile_table(proc)
    k_ms, k_launches, k_bytes = prof["arw2_decode"]
    avg = k_ms / max(k_launches, 1)
    achieved = k_bytes / (avg * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        t0 = time.perf_counter()
        st, _ = oracle.decode_arw2(strip, spec.width, spec.height, synth.SONY_CURVE)
        dt = time.perf_counter() - t0
        cpu = {"value": round(mp / dt, 2), "unit": "MP/s", "cores": 1, "kind": "port",
               "sample": f"the whole {spec.width}x{spec.height} strip, {dt:.2f} s, status {st}"}
    proc.close()
    line = {
        "metric": "RAW ingest MP/s: Sony compressed ARW -> resident u16 mosaic",
        "value": round(mp / ingest_s, 1), "unit": "MP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ingest_s * 1e3, 3), "higher_is_better": True,
        "dtype": "u16", "data": "synthetic",
        "decode_kernel_ms": round(float(np.median(gpu_ms)), 4),
        "u16_upload_ms_per_step": round(upload_s * 1e3, 3),
        "roofline": {"bound": "hbm", "kernel": "arw2_decode", "avg_launch_ms": round(avg, 4),
                     "algorithmic_bytes_per_launch": int(k_bytes), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4)},
        "cpu_baseline": cpu,
        "config": {"workload": "6048x4024 RGGB 14-bit mosaic (seed 0x52480002) as Sony "
                               "compressed ARW, 24.3 MB strip in a TIFF container, decoded "
                               "from host memory", "parallelism": "single GPU"},
    }
    print(json.dumps(line))


def detail_cpu_baseline(rgb, detail):
    from oracle import oracle
    h, w = rgb.shape[:2]
    t0 = time.perf_counter()
    st, _, _ = oracle.detail_stage(
        rgb, false_colour=detail.false_colour, chroma=detail.nr_chroma,
        luminance=detail.nr_luminance, luminance_detail=detail.nr_luminance_detail,
        balance=detail.balance[:], ceilings=detail.ceilings[:],
        capture=(detail.amount, detail.sigma, detail.noise_protection, detail.iterations))
    dt = time.perf_counter() - t0
    return {"value": round(w * h / 1e6 / dt, 3), "unit": "MP/s", "cores": 1, "kind": "port",
            "sample": f"the detail stage alone on a {w}x{h} crop of the demosaiced frame, "
                      f"{dt:.2f} s, status {st}"}


def profile_table(proc):
    """{kernel group: (total ms, launches, algorithmic bytes per launch)}."""
    import ctypes
    lib = proc.lib
    out = {}
    n = lib.rhd_profile_count(proc.ctx)
    for k in range(n):
        name = ctypes.c_char_p()
        ms = ctypes.c_double()
        launches = ctypes.c_uint64()
        nbytes = ctypes.c_uint64()
        lib.rhd_profile_entry(proc.ctx, k, ctypes.byref(name), ctypes.byref(ms),
                              ctypes.byref(launches), ctypes.byref(nbytes))
        out[name.value.decode()] = (ms.value, launches.value, nbytes.value)
    return out


if __name__ == "__main__":
    main()
