"""Benchmark: Mpaths/s and Mrays/s of the HIP wavefront path tracer on one GPU.

Workload (the configuration the metric is quoted on): one displaced
icosphere of 1,310,720 triangles (level-8 subdivision, 4-octave value noise,
amplitude 5%) with a GGX conductor (roughness 0.3) on a diffuse ground quad,
sun + gradient sky, 1920x1080, max depth 8 with Russian roulette from depth
3, FP32, per-pixel seeds of the reference's CPU integrator. Synthetic
geometry from a fixed seed; all scene data is resident in HBM before timing.

A step renders `--spp-per-step` samples of every pixel (default 256 = the
whole configuration) through the full wavefront loop; W warmup steps, then
K timed steps bracketed by device synchronisation. `value` = pixels x samples
/ seconds / 1e6 over the timed steps.

roofline: the closest-hit traversal kernel. Its algorithmic bytes per ray are
32 B ray in (bounce rays only: camera rays are generated in the kernel from
their queue index) + 16 B hit out, plus 32 B per BVH node box tested, 36 B
per triangle test and 96 B per instance visit; achieved = those bytes / the
kernel's summed launch time measured with HIP events on the library's stream
during the timed steps. The timed steps run the production kernels, which
keep no work counters; the node / triangle / instance counts come from
rendering the same samples again (untimed) through the counting variants of
the traversal kernels, which trace exactly the same rays.

guided: the guided-vs-plain comparison at equal wall-clock time on the
caustic room at 1280x720 (a 499,280-triangle delta-glass sphere in a closed
diffuse room lit only through a wall opening): a plain render of 16384
samples (sample indices disjoint from the compared runs) is the reference,
its own noise estimated from its two halves and subtracted; plain and guided
renders each get the same wall-clock budget (guided: training after every
batch included); relMSE = mean over finite pixels of the channel mean of
(x - r)^2 / (r^2 + 1e-2), also reported without the 0.1 % largest pixel
errors, and the median pixel error.

cpu_baseline: the reference's CPU tiled integrator, compiled from its sources
with its Release flags (oracle/_ref) where present, else the oracle's
restatement with the same flags; timed on this host over a bounded sample of
the same workload (whole frames at 1 sample each, sized to ~10 s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + Mrays/s (1 GPU); relMSE at equal time, guided vs plain"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(st: dict, kind: str = "closest", camera_rays: int = 0) -> float:
    if kind == "closest":
        return (48.0 * st["closest_rays"] - 32.0 * camera_rays + 32.0 * st["node_tests"] + 36.0 * st["tri_tests"]
                + 96.0 * st["inst_visits"])
    return (36.0 * st["shadow_rays"] + 32.0 * st["shadow_node_tests"] + 36.0 * st["shadow_tri_tests"]
            + 96.0 * st["shadow_inst_visits"])


def read_traffic():
    """HBM bytes per closest-hit launch from the committed rocprofv3 counter
    summary, if one exists (profiles/closest_hit_traffic.txt)."""
    path = os.path.join(ROOT, "profiles", "closest_hit_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0].startswith("bytes_per_launch"):
            try:
                vals[parts[0]] = float(parts[1])
            except ValueError:
                pass
    return vals.get("bytes_per_launch")


def relmse(x, r):
    import numpy as np
    ok = np.isfinite(x) & np.isfinite(r)
    return float(np.mean((x[ok] - r[ok]) ** 2 / (r[ok] ** 2 + 1e-2)))


def relmse_stats(x, r, ref_err=None, mask=None):
    """relMSE of image x against reference r (mean over finite pixels of the
    channel-mean (x - r)^2 / (r^2 + 1e-2)), the same with the 0.1 % largest
    pixel errors left out (single-path outliers dominate the mean at short
    times), and the median pixel error. ref_err: per-pixel error the
    reference's own noise adds (subtracted, see guided_comparison); mask:
    only these pixels."""
    import numpy as np
    e = (x - r) ** 2 / (r ** 2 + 1e-2)
    e = e.reshape(-1, 3).mean(1)
    if ref_err is not None:
        e = e - ref_err
    if mask is not None:
        e = e[mask]
    e = e[np.isfinite(e)]
    s = np.sort(e)
    keep = s[: max(1, int(len(s) * 0.999))]
    return {"relmse": round(float(e.mean()), 6), "relmse_trim999": round(float(keep.mean()), 6),
            "median": round(float(np.median(e)), 6)}


def guided_comparison(seconds: float, width: int, height: int, ref_spp: int, freq: int):
    """Plain vs guided relMSE at equal wall-clock on the caustic room.

    Reference: a plain render of ref_spp samples from sample indices disjoint
    from the compared runs, made of two halves; its own noise adds about a
    quarter of the halves' mutual error to every measured pixel error (the
    difference of two halves has four times the variance of the full
    reference), and that estimate is subtracted. The guide uses the
    library's defaults (16-level hash grid of 2^19 entries x 2 FP16
    features, 3x64 MLP, 8 vMF lobes; after every batch one training step
    per 2^22 of its paths, on slices of up to 2^18 samples drawn from it).
    The plain run renders in the library's default batches (about 2^27
    paths), the guided run in whole guided batches (the library caps them at
    2^26 paths, so that the guide learns often); the guided run's time
    includes its training."""
    import numpy as np
    from xenon_amd import HipRenderer, HipScene, scenes

    desc = scenes.caustic_room(width=width, height=height, freq=freq)
    hs = HipScene(desc)
    r = HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces)
    r.upload_scene(hs)
    cam = hs.camera()
    t0 = time.perf_counter()
    halves = []
    for h in range(2):
        r.reset()
        done = 0
        half = ref_spp // 2
        while done < half:
            n = min(256, half - done)
            r.render(cam, 10_000_001 + h * half + done, n, want_image=False)
            done += n
        acc, _ = r.read_accum()
        halves.append(acc / half)
    ref = 0.5 * (halves[0] + halves[1])
    d = (halves[0] - halves[1]) ** 2 / (ref ** 2 + 1e-2)
    ref_err = 0.25 * d.reshape(-1, 3).mean(1)
    log(f"guided: reference {ref_spp} spp in {time.perf_counter() - t0:.1f} s")

    def run(guided: bool):
        rr = r
        rr.reset()
        if guided:
            rr.guide_enable()
        chunk = max(rr.tuning()["spp_per_launch"], 1)
        if guided:   # whole guided batches (the library caps them at 2^26 paths)
            chunk = max(1, min(chunk, (1 << 26) // (desc.width * desc.height)))
        t = time.perf_counter()
        spp = 0
        while time.perf_counter() - t < seconds:
            rr.render(cam, 1 + spp, chunk, want_image=False)
            spp += chunk
        el = time.perf_counter() - t
        img, n = rr.read_accum()
        if guided:
            rr.guide_disable()
        return img / n, n, el

    plain, sp, tp = run(False)
    guided, sg, tg = run(True)
    ep, eg = relmse_stats(plain, ref, ref_err), relmse_stats(guided, ref, ref_err)
    log(f"guided: plain {sp} spp {ep}; guided {sg} spp {eg}")
    r.close()
    # the caustic region: pixels the glass makes at least 1.5x brighter than
    # a diffuse white sphere in its place does (4096 plain samples of that
    # variant); the errors are split between it and the rest of the room
    dd = scenes.caustic_room(width=width, height=height, freq=freq)
    dd.materials[3] = scenes.Material(base=(0.75, 0.75, 0.75), roughness=1.0, ior=1.0)
    hd = HipScene(dd)
    with HipRenderer(dd.width, dd.height, dd.max_bounces, dd.min_bounces) as r2:
        r2.upload_scene(hd)
        r2.render(hd.camera(), 20_000_001, 4096, want_image=False)
        a2, n2 = r2.read_accum()
    lum = np.array([0.2126, 0.7152, 0.0722], np.float32)
    caustic = (ref @ lum) > 1.5 * ((a2 / n2) @ lum) + 1e-3

    # runs stop at the first batch boundary past the budget, so their
    # wall-clock differs a little; relMSE x seconds (the usual efficiency)
    # corrects for that, since relMSE falls as 1/samples
    def ratio(a, b, k):
        return round(a[k] * tp / (b[k] * tg), 3) if b[k] > 0 else None

    regions = {}
    for name, m in (("caustic", caustic), ("rest", ~caustic)):
        pa, ga = relmse_stats(plain, ref, ref_err, m), relmse_stats(guided, ref, ref_err, m)
        regions[name] = {"pixels": int(m.sum()), "trim999_ratio_plain_over_guided": ratio(pa, ga, "relmse_trim999"),
                         "median_ratio_plain_over_guided": ratio(pa, ga, "median")}
    return {"workload": f"caustic room ({desc.num_tris} tris), {desc.width}x{desc.height}, depth {desc.max_bounces}",
            "seconds": round(seconds, 2), "reference_spp": ref_spp,
            "plain": {"spp": sp, "seconds": round(tp, 2), **ep},
            "guided": {"spp": sg, "seconds": round(tg, 2), **eg},
            "relmse_ratio_plain_over_guided": ratio(ep, eg, "relmse"),
            "trim999_ratio_plain_over_guided": ratio(ep, eg, "relmse_trim999"),
            "median_ratio_plain_over_guided": ratio(ep, eg, "median"),
            "regions": regions}


class _StdoutToStderr:
    """Sends what native code writes to stdout (the reference's build prints
    progress lines) to stderr, so that stdout holds only the JSON line."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        import ctypes
        ctypes.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)


def cpu_baseline(desc, budget_s: float):
    """Times the CPU integrator on this host; returns the cpu_baseline object."""
    with _StdoutToStderr():
        return _cpu_baseline(desc, budget_s)


def _cpu_baseline(desc, budget_s: float):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    kind = "reference" if os.path.exists(pyoracle.PATHS["ref_fast"]) else "port"
    lib = pyoracle.Lib("ref_fast" if kind == "reference" else "fast")
    osc = lib.load(desc)
    W, H = desc.width, desc.height
    threads = os.cpu_count() if kind == "reference" else pyoracle.default_threads()
    # one frame first, then as many as fit the budget (at least one more)
    t0 = time.perf_counter()
    osc.render_average(1) if kind == "reference" else osc.render_sum(1, 1, threads)
    one = time.perf_counter() - t0
    n = max(1, min(512, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    if kind == "reference":
        osc.render_average(n)
    else:
        osc.render_sum(1, n, threads)
    dt = time.perf_counter() - t0
    return {"value": round(W * H * n / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": kind,
            "cpu_model": cpu_model(),
            "sample": f"{n} frame(s) of the same scene at {W}x{H}, 1 sample per pixel each ({W * H * n} paths, "
                      f"{dt:.1f} s), max depth {desc.max_bounces}; paths are independent so Mpaths/s is "
                      f"per-path throughput"}


def simd_use(tot: dict) -> dict:
    """Active-lane fractions of the traversal phases, from the counting
    re-run: lane steps over 64 x wave iterations of the interior phase (node
    pairs), the leaf phase (held leaves) and the round loop (lanes holding a
    ray), with the node pairs and leaf steps per ray."""
    def frac(a, b):
        return round(tot[a] / (64.0 * tot[b]), 4) if tot.get(b) else None

    out = {}
    for pre, rays in (("", "closest_rays"), ("shadow_", "shadow_rays")):
        out[(pre or "closest_") + "interior"] = frac(pre + "interior", pre + "interior_iters")
        out[(pre or "closest_") + "leaf"] = frac(pre + "leaf_steps", pre + "leaf_iters")
        out[(pre or "closest_") + "round"] = frac(pre + "round_lanes", pre + "round_iters")
        if tot.get(rays):
            out[(pre or "closest_") + "pairs_per_ray"] = round(tot[pre + "interior"] / tot[rays], 3)
            out[(pre or "closest_") + "leaf_steps_per_ray"] = round(tot[pre + "leaf_steps"] / tot[rays], 3)
            out[(pre or "closest_") + "wave_iters_per_ray"] = round(
                (tot[pre + "interior_iters"] + tot[pre + "leaf_iters"]) * 64.0 / tot[rays], 3)
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


CONFIG_NAMES = {0: "cornell", 1: "big_mesh", 2: "instancing", 3: "batched_jobs"}


class Workload:
    """One configuration set up for timing: `step(k)` renders step k's
    samples (distinct sample indices per step), `paths` per step."""

    def __init__(self, name, renderer, step, paths, config, data, cpu_descs, extra=None):
        self.name, self.r, self.step, self.paths = name, renderer, step, paths
        self.config, self.data, self.cpu_descs, self.extra = config, data, cpu_descs, extra or {}


def setup_single(args, desc, label):
    """Configurations 0-2: one scene, `sps` samples of every pixel per step."""
    from xenon_amd import HipRenderer, HipScene

    t0 = time.perf_counter()
    hs = HipScene(desc)
    log(f"scene: {desc.num_tris} triangles, built in {time.perf_counter() - t0:.1f} s")
    r = HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces,
                    spp_per_launch=args.spp_per_launch, block_size=args.block, lds_stack=args.lds_stack)
    t0 = time.perf_counter()
    r.upload_scene(hs)
    cam = hs.camera()
    log(f"upload {time.perf_counter() - t0:.2f} s")
    sps = args.spp_per_step or desc.spp

    def step(k):
        r.render(cam, 1 + k * sps, sps, want_image=False)

    cfg = {"workload": f"{label}, {desc.width}x{desc.height}, {sps} spp per step, depth {desc.max_bounces}, "
                       f"RR from {desc.min_bounces}",
           "width": desc.width, "height": desc.height, "spp_per_step": sps, "triangles": desc.num_tris}
    return Workload(label, r, step, desc.width * desc.height * sps, cfg, "synthetic (seeded scene generator)",
                    [(desc, 1.0)])


def setup_jobs(args):
    """Configuration 3: the 32 jobs of the seeded job list (seed 6), each on
    the full-size scene of its generator (Cornell, the L8 displaced
    icosphere, the 4096-instance scene; one device-resident copy of each,
    shared by the jobs of that generator, each job with its own camera
    aspect), all rendered by one xnh_render_jobs call per step."""
    from xenon_amd import HipRenderer, HipScene, scenes
    from xenon_amd.renderer import GpuScene

    jobs = scenes.job_list()
    gens = {0: scenes.cornell(512, 512), 1: scenes.big_mesh(), 2: scenes.instancing()}
    t0 = time.perf_counter()
    hss = {k: HipScene(d) for k, d in gens.items()}
    log(f"job scenes built in {time.perf_counter() - t0:.1f} s")
    r = HipRenderer(1920, 1080, 12, 3, spp_per_launch=args.spp_per_launch, block_size=args.block,
                    lds_stack=args.lds_stack)
    r.upload_scene(hss[1])
    tuned = None
    if args.autotune:
        t0 = time.perf_counter()
        tuned = r.autotune(hss[1].camera(), cache=None)
        log(f"autotune {time.perf_counter() - t0:.1f} s: {tuned}")
    gs = {k: GpuScene(r, h) for k, h in hss.items()}
    spec = []
    for j in jobs:
        d = gens[j.kind]
        spec.append(dict(scene=gs[j.kind], camera=hss[j.kind].camera(j.width, j.height), width=j.width,
                         height=j.height, spp_count=j.spp, max_bounces=j.max_bounces, min_bounces=d.min_bounces))
    paths = sum(j.width * j.height * j.spp for j in jobs)

    def step(k):
        for sj, j in zip(spec, jobs):
            sj["spp_first"] = 1 + k * j.spp
        r.render_jobs(spec, want_images=False)

    kinds = [sum(1 for j in jobs if j.kind == g) for g in range(3)]
    cfg = {"workload": f"{len(jobs)} batched jobs (seeded list, seed 6): {kinds[0]} Cornell, {kinds[1]} displaced "
                       f"icosphere L8 ({gens[1].num_tris} tris), {kinds[2]} instancing (4096 instances, "
                       f"{gens[2].num_tris} unique tris); resolutions 256^2-2048^2, spp 16-1024, depth 4-12; one "
                       f"render_jobs call per step",
           "jobs": len(jobs), "paths_per_step": paths,
           "tuning": tuned if tuned is not None else r.tuning()}
    # CPU sample: each generator's scene at the resolution of its first job,
    # weighted by that generator's share of the job list's paths
    w = [sum(j.width * j.height * j.spp for j in jobs if j.kind == g) / paths for g in range(3)]
    cpu = []
    for g in range(3):
        first = next((j for j in jobs if j.kind == g), None)
        if first is None:
            continue
        d = {0: scenes.cornell, 1: scenes.big_mesh, 2: scenes.instancing}[g](width=first.width, height=first.height,
                                                                           max_bounces=first.max_bounces)
        cpu.append((d, w[g]))
    return Workload("batched_jobs", r, step, paths, cfg, "synthetic (seeded scene generators and job list)", cpu,
                    extra={"gpu_scenes": gs})


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 3],
                    help="0 Cornell 512^2 64 spp; 1 displaced icosphere 1080p 256 spp (the default line); "
                         "2 instancing 1080p 64 spp; 3 the 32 batched jobs")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 3; 10 for config 0; 1 for 3)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=0, help="0: the configuration's spp")
    ap.add_argument("--level", type=int, default=8, help="icosphere subdivision level (8 = the benchmark)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp-per-launch", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lds-stack", type=int, default=0)
    ap.add_argument("--autotune", type=int, default=1, help="config 3: run the autotuner first (1) or not (0)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--guided-seconds", type=float, default=8.0,
                    help="config 1: wall-clock budget of each side of the guided-vs-plain comparison (0: skip)")
    ap.add_argument("--guided-width", type=int, default=1280)
    ap.add_argument("--guided-height", type=int, default=720)
    ap.add_argument("--guided-ref-spp", type=int, default=16384)
    ap.add_argument("--guided-freq", type=int, default=158)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the reference kernel "
                         "for scaling HBM counter readings) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the renderer shards by "
              f"samples across GPUs by design but this build runs on one device).", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {0: 10, 1: 3, 2: 3, 3: 1}[args.config]

    import torch

    if args.copy_calibration:
        n = args.copy_calibration << 18   # float32 elements
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                          "launches": 5}), flush=True)
        return

    from xenon_amd import scenes

    if args.config == 0:
        wl = setup_single(args, scenes.cornell(512, 512, 64), "Cornell box (36 tris, diffuse, area light)")
    elif args.config == 1:
        wl = setup_single(args, scenes.big_mesh(width=args.width, height=args.height, level=args.level),
                          f"displaced icosphere L{args.level} + ground, GGX conductor r=0.3, sun+sky")
    elif args.config == 2:
        desc = scenes.instancing()
        wl = setup_single(args, desc, f"instancing: TLAS over 4096 instances of 10 displaced-icosphere BLAS "
                                      f"({desc.num_tris} unique tris), 40/30/30 diffuse/GGX/dielectric, "
                                      f"64 emissive tris")
    else:
        wl = setup_jobs(args)
    r = wl.r
    for w in range(args.warmup):
        wl.step(args.steps + w)   # sample indices after the timed steps'
        log(f"warmup step {w + 1}/{args.warmup}: {r.stats()['render_ms']:.1f} ms")
    r.set_profiling(timing=True)
    keys = ["closest_rays", "shadow_rays", "nonfinite", "closest_ms", "shadow_ms", "shade_ms", "closest_launches",
            "render_ms"]
    work_keys = ["node_tests", "interior", "tri_tests", "inst_visits", "shadow_node_tests", "shadow_interior",
                 "shadow_tri_tests", "shadow_inst_visits", "interior_iters", "leaf_iters", "leaf_steps",
                 "round_iters", "round_lanes", "shadow_interior_iters", "shadow_leaf_iters", "shadow_leaf_steps",
                 "shadow_round_iters", "shadow_round_lanes"]
    tot = {k: 0 for k in keys + work_keys}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        wl.step(k)
        st = r.stats()
        for key in keys:
            tot[key] += st[key]
        log(f"step {k + 1}/{args.steps}: {st['render_ms']:.1f} ms device")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    # Work counts of the timed steps: the same samples again through the
    # counting kernel variants (every path is a function of its (pixel,
    # sample) seed, so the rays, and the nodes and triangles they test, are
    # identical); untimed.
    r.set_profiling(work=True)
    for k in range(args.steps):
        wl.step(k)
        st = r.stats()
        for key in work_keys:
            tot[key] += st[key]
    r.set_profiling()
    paths = wl.paths * args.steps
    mpaths = paths / elapsed / 1e6
    mrays = (tot["closest_rays"] + tot["shadow_rays"]) / elapsed / 1e6
    camera_rays = wl.paths * args.steps   # one per path
    cbytes = algorithmic_bytes(tot, "closest", camera_rays)
    launches = max(1, tot["closest_launches"])
    achieved = cbytes / (tot["closest_ms"] / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic() if args.config == 1 else None,
                "kernel": "k_closest (closest-hit traversal)",
                "bytes_per_launch": round(cbytes / launches), "avg_launch_ms": round(tot["closest_ms"] / launches, 4)}
    sbytes = algorithmic_bytes(tot, "shadow")
    # shading: path state (64 B; not for camera paths, which are rebuilt
    # from their index) and hit (20 B) in, material record (64 B), radiance
    # term (16 B) out per shaded entry; the continuation (64 B) of every
    # closest-hit ray after the camera rays; one shadow record (48 B) per
    # shadow ray
    bounce = max(0, tot["closest_rays"] - camera_rays)
    hbytes = 100.0 * tot["closest_rays"] + (64.0 + 64.0) * bounce + 48.0 * tot["shadow_rays"]
    out = {
        "metric": METRIC,
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": wl.data,
        "config": {"config": args.config, "name": CONFIG_NAMES[args.config], **wl.config},
        "mrays_per_s": round(mrays, 3),
        "rays_per_path": round((tot["closest_rays"] + tot["shadow_rays"]) / paths, 3),
        "kernel_ms": {"closest": round(tot["closest_ms"], 2), "shade": round(tot["shade_ms"], 2),
                      "shadow": round(tot["shadow_ms"], 2), "device_total": round(tot["render_ms"], 2)},
        "shadow_roofline": {"achieved": round(sbytes / (tot["shadow_ms"] / 1e3) / 1e9, 2) if tot["shadow_ms"] else None,
                            "frac": round(sbytes / (tot["shadow_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                            if tot["shadow_ms"] else None, "unit": "GB/s", "kernel": "k_shadow (any-hit traversal)"},
        "shade_roofline": {"achieved": round(hbytes / (tot["shade_ms"] / 1e3) / 1e9, 2) if tot["shade_ms"] else None,
                           "frac": round(hbytes / (tot["shade_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                           if tot["shade_ms"] else None, "unit": "GB/s", "kernel": "k_shade (shading)"},
        "rays": {"paths": wl.paths * args.steps, "closest": int(tot["closest_rays"]), "shadow": int(tot["shadow_rays"])},
        "nonfinite_paths": int(tot["nonfinite"]),
        "simd": simd_use(tot),
        "work_per_ray": {k: round(tot[k] / max(1, tot[r]), 3) for k, r in (
            ("node_tests", "closest_rays"), ("tri_tests", "closest_rays"), ("inst_visits", "closest_rays"),
            ("shadow_node_tests", "shadow_rays"), ("shadow_tri_tests", "shadow_rays"),
            ("shadow_inst_visits", "shadow_rays"))},
        "roofline": roofline,
    }
    for g in wl.extra.get("gpu_scenes", {}).values():
        g.release()
    r.close()
    if args.config == 1 and args.guided_seconds > 0:
        out["guided"] = guided_comparison(args.guided_seconds, args.guided_width, args.guided_height,
                                          args.guided_ref_spp, args.guided_freq)
    if not args.no_cpu:
        log("timing the CPU integrator ...")
        cb = cpu_baseline_weighted(wl.cpu_descs, args.cpu_seconds)
        cb["mpaths_ratio_gpu_over_cpu"] = round(mpaths / cb["value"], 1) if cb["value"] else None
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def cpu_baseline_weighted(descs, budget_s: float):
    """The CPU baseline of one scene, or of several scenes weighted by their
    share of the workload's paths (batched jobs): the weighted harmonic mean
    of their rates, i.e. total paths / total CPU time."""
    if len(descs) == 1:
        return cpu_baseline(descs[0][0], budget_s)
    parts = [(cpu_baseline(d, budget_s / len(descs)), w) for d, w in descs]
    wsum = sum(w for _, w in parts)
    rate = wsum / sum(w / p["value"] for p, w in parts)
    first = parts[0][0]
    return {"value": round(rate, 4), "unit": "Mpaths/s", "cores": first["cores"], "kind": first["kind"],
            "cpu_model": first["cpu_model"],
            "sample": "per generator, the scene of its first job at that job's resolution and depth, timed as: " +
                      " | ".join(f"weight {w:.3f}: {p['value']} Mpaths/s over {p['sample']}" for p, w in parts) +
                      "; combined as total paths over total CPU time"}


if __name__ == "__main__":
    main()
