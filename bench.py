#!/usr/bin/env python3
"""Training-step throughput of the TurboNeRF hot path on one MI355X.

Workload (the reference's full training step, src/core/trainer.cu:313-349 and
src/core/nerf-network.cu:547-781): hash grid 16 levels x 2 features, 2^19
entries per level, resolution 16 -> 2048, table init U(+-1e-4); density MLP
32 -> 64 -> 16, colour MLP (16 + 16 SH) -> 64 -> 64 -> 3; 2^18 compacted
samples per step; targets are an 8-blob Gaussian field rendered at 800x800
from 100 random cameras on a radius-2 sphere (fov 50 deg), composited over a
random background; smooth-L1 loss; Adam lr 1e-2, betas (0.9, 0.99), eps 1e-15.

Per step: ray generation, occupancy-grid marching (count, scan/select,
generate), hash-grid forward, fused MLP forward, compositing + loss +
backward, fused MLP backward, hash-grid scatter-add, Adam. The occupancy grid
is refreshed every 16 steps during warm-up and timed on its own.

value = samples trained per second (device-side sample counter over the timed
steps) for fp16-MFMA mixed precision; the plain fp32 path, the per-phase
breakdown, the occupancy update and render throughput go under "extra".
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

BATCH = 1 << 18
N_CAMS, RES = 100, 800
SCENE_SEED = 7

# algorithmic bytes (B) or FLOPs per unit for each trainer phase:
# (bound, per-sample, per-ray, per-parameter)
ALGO = {
    "rays": ("hbm", 0, 60, 0),
    "march_count": ("hbm", 0, 38, 0),
    "scan_select": ("hbm", 0, 8, 0),
    "march_generate": ("hbm", 44, 38, 0),
    # positions 12 B + 16 levels x 8 corners x the table's feature pair
    # (4 B fp16, 8 B fp32) + 32 fp32 features out (128 B)
    "hashgrid_forward": ("hbm", 12 + 512 + 128, 0, 0),
    "mlp_forward": ("mfma", 18816, 0, 0),
    "composite": ("hbm", 34, 60, 0),
    "distortion": ("hbm", 24, 8, 0),
    "mlp_backward": ("mfma", 56448 - 18816, 0, 0),
    # positions 12 B + 32 fp32 feature gradients 128 B + 16 levels x 8
    # corners x the fp32 gradient pair read and written (8 B)
    "hashgrid_backward": ("hbm", 12 + 128 + 1024, 0, 0),
    "adam": ("hbm", 0, 0, 36),
    "loss_reduce": ("hbm", 0, 16, 0),
}
# with Adam on the hash table fused into the hash-grid backward's sum passes
# (the default): the backward moves the per-sample inputs and, per table
# parameter, the Adam state (fp32 weight, moments and step read and written,
# the fp16 copy written: 34 B; 32 B in fp32 training) instead of gradient
# pairs per corner; Adam's own pass covers the network weights (36 B each)
ALGO_FUSED = {
    "hashgrid_backward": ("hbm", 12 + 128, 0, 34),
    "adam": ("hbm", 0, 0, 36),
}
# phases whose bytes depend on the network precision (fp32 tables)
ALGO_FP32 = {"hashgrid_forward": ("hbm", 12 + 1024 + 128, 0, 0)}
# the kernels each phase launches (rocprofv3 kernel names)
KERNEL = {
    "hashgrid_forward": ("hashgrid_fwd_kernel",),
    # the plan (count, reduce, scan) is formed with the batch on the
    # generation stream; the network streams' phase runs the gradient passes
    "hashgrid_backward": ("hashgrid_bwd_scatter_kernel", "hashgrid_bwd_accumulate_kernel",
                          "hashgrid_bwd_finish_kernel"),
    "hashgrid_plan": ("hashgrid_bwd_count_kernel", "hashgrid_bwd_reduce_kernel", "hashgrid_bwd_scan_kernel"),
    "mlp_forward": ("mlp_forward_kernel",),
    "mlp_backward": ("mlp_backward_kernel", "reduce_partials_kernel"),
    "composite": ("composite_chunk_kernel", "composite_ray_kernel", "composite_grad_kernel"),
    "loss_reduce": ("step_metrics_kernel",),
    "march_count": ("march_count_kernel",),
    "march_generate": ("march_generate_kernel", "march_expand_kernel"),
    "adam": ("adam_kernel",),
}
# phases on the network stream: the step's critical path (batch generation
# runs on its own stream, overlapped with them)
NETWORK_PHASES = ("hashgrid_forward", "mlp_forward", "distortion", "composite", "mlp_backward", "hashgrid_backward",
                  "adam", "loss_reduce")
METRIC = "train samples/sec (2^18/step) mixed vs FP32 at \u22641e-3 grad err; render Mrays/s; %HBM/MFMA peak"
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp16": 2500.0, "fp32": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed training steps")
    ap.add_argument("--warmup", type=int, default=2048, help="untimed steps (occupancy update every 16)")
    ap.add_argument("--precision", choices=["mixed", "fp32", "both"], default="both")
    ap.add_argument("--cpu-samples", type=int, default=1 << 18, help="samples per CPU-baseline step")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--no-extra-configs", action="store_true",
                    help="skip the distortion-loss and eight-scene training measurements")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="profiling run: render the targets, then only --warmup and this many mixed-precision "
                         "training steps (no occupancy updates, second pass, render or CPU leg), and exit")
    ap.add_argument("--summarize", nargs="+", metavar="CSV",
                    help="summarise rocprofv3 CSVs of a --profile-steps run: kernel_trace.csv [FETCH_SIZE "
                         "counter_collection.csv WRITE_SIZE counter_collection.csv copy-probe FETCH_SIZE csv "
                         "copy-probe WRITE_SIZE csv]; prints markdown (no GPU needed)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="summarise a --profile-steps run's kernel_trace.csv and counter_collection.csv files "
                         "(any counters) for the MLP kernels; prints markdown (no GPU needed)")
    ap.add_argument("--profile-render", type=int, default=0, metavar="FRAMES",
                    help="profiling run: --warmup mixed-precision training steps (occupancy update every 16), "
                         "then FRAMES renders of an 800x800 frame of the trained NeRF, and exit")
    ap.add_argument("--summarize-after", default="write_rgba_kernel", metavar="KERNEL",
                    help="summarise the dispatches after the last dispatch of KERNEL (adam_kernel for a "
                         "--profile-render run)")
    ap.add_argument("--summarize-units", type=int, default=0, metavar="N",
                    help="divide totals by N (frames of a --profile-render run) instead of the Adam steps")
    ap.add_argument("--summarize-last", type=int, default=0,
                    help="with --summarize: only the last N training steps (after the warm-up)")
    ap.add_argument("--samples-per-step", type=float, default=0.0,
                    help="with --summarize: samples per step, for the per-sample traffic column")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB 10 times (calibrates HBM counters)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain copy kernel of known size, for scaling rocprofv3's FETCH_SIZE /
    WRITE_SIZE readings to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_launch": 2 * n * 4, "read_bytes": n * 4, "write_bytes": n * 4}))


def make_scene():
    import numpy as np

    from turbonerf_amd import synthetic

    t0 = time.perf_counter()
    blobs, cams, images = synthetic.blob_dataset(seed=SCENE_SEED, n_cams=N_CAMS, width=RES, height=RES)
    log(f"[bench] targets: {N_CAMS} x {RES}x{RES} rendered in {time.perf_counter() - t0:.1f} s, "
        f"alpha mean {images[..., 3].mean() / 255:.3f}")
    return blobs, cams, np.ascontiguousarray(images)


def make_trainer(tn, synthetic, scene, precision):
    blobs, cams, images = scene
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    ds = tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images)
    nerf.attach_dataset(ds)
    cfg = tn.TrainingConfig()
    cfg.network.precision = precision
    cfg.network.beta1 = 0.9
    cfg.network.beta2 = 0.99
    cfg.network.epsilon = 1e-15
    cfg.network.learning_rate = 1e-2
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=BATCH)
    tr.load_images()
    return mgr, nerf, tr


def run_precision(tn, synthetic, scene, precision, args):
    import torch

    mgr, nerf, tr = make_trainer(tn, synthetic, scene, precision)
    name = "mixed" if precision == 1 else "fp32"
    # warm-up: the reference's loop, occupancy grid refreshed every 16 steps
    step = 0
    while step < args.warmup:
        k = min(16, args.warmup - step)
        tr.train_steps(k, read_metrics=False, graph=True)
        step += k
        if step % 16 == 0:
            om = tr.update_occupancy_grid(step)
            if step % 512 == 0:
                log(f"[bench {name}] warm-up step {step}: occupied {om.n_occupied / om.n_total:.3f}")
    # launch shapes from the autotuner, timed on this workload (its trial
    # rounds are training steps too, part of the warm-up)
    tune = tr.autotune(rounds_per_trial=16, repetitions=3)
    step = nerf.training_step
    log(f"[bench {name}] autotuned launch shapes {tune['picks']}: {tune['ns_per_sample_before']:.3f} -> "
        f"{tune['ns_per_sample_after']:.3f} ns/sample; trials (value, ns/sample): "
        + "; ".join(f"{k} " + " ".join(f"{v}:{t:.3f}" for v, t in rows) for k, rows in tune["trials"].items()))
    m = tr.train_step()
    step += 1
    log(f"[bench {name}] warm-up done: loss {m.loss:.4f} rays {m.n_rays} samples {m.n_samples}")

    # timed region: K steps, each the replay of the captured training-step
    # graph (captured during the warm-up), no host synchronisation inside
    tr.set_timing(False)
    tr.train_steps(1, read_metrics=False, graph=True)  # graphs captured before the clock starts
    step += 1

    def timed(graph):
        s0 = tr.samples_trained()
        tr.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr.train_steps(args.steps, read_metrics=False, graph=graph)
        tr.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return tr.samples_trained() - s0, wall

    samples, wall = timed(True)
    step += args.steps
    res = dict(samples_per_s=samples / wall, ms_per_step=1e3 * wall / args.steps,
               samples_per_step=samples / args.steps)
    # the same K steps enqueued kernel by kernel, for comparison
    samples_e, wall_e = timed(False)
    step += args.steps
    res["eager_samples_per_s"] = samples_e / wall_e
    res["eager_ms_per_step"] = 1e3 * wall_e / args.steps

    # the same K steps again with HIP events around every phase on the
    # trainer's stream (per-kernel durations for the roofline)
    tr.reset_timings()
    tr.set_timing(True)
    tr.train_steps(args.steps, read_metrics=False)
    tr.synchronize()
    phases = tr.phase_timings()
    tr.set_timing(False)
    step += args.steps
    rays = []
    for _ in range(8):
        m = tr.train_step()
        rays.append(m.n_rays)
        step += 1
    res["rays_per_step"] = statistics.mean(rays)
    res["loss"] = m.loss
    res["phases"] = fold_phases(phases)

    # occupancy update, timed on its own
    tr.synchronize()
    t0 = time.perf_counter()
    n_upd = 8
    for i in range(n_upd):
        om = tr.update_occupancy_grid(step + 16 * (i + 1))
    tr.synchronize()
    res["occupancy_update_ms"] = 1e3 * (time.perf_counter() - t0) / n_upd
    res["occupied_fraction"] = om.n_occupied / om.n_total
    res["bitfield"] = tr.get_occupancy_bitfield()
    res["trainer"] = (mgr, nerf, tr)
    res["autotune"] = dict(picks=tune["picks"], ns_per_sample_before=round(tune["ns_per_sample_before"], 4),
                           ns_per_sample_after=round(tune["ns_per_sample_after"], 4))
    log(f"[bench {name}] {res['samples_per_s'] / 1e6:.2f} M samples/s, {res['ms_per_step']:.3f} ms/step "
        f"(captured step; eager {res['eager_samples_per_s'] / 1e6:.2f} M samples/s), "
        f"{res['samples_per_step']:.0f} samples/step, {res['rays_per_step']:.0f} rays/step")
    return res


def fold_phases(phases):
    """ms per step of each phase from the trainer's timings (name -> (total
    ms, launches)). A step over the batch's two halves times the second
    half's phases as "<phase>/2" on its own stream and each half's hash-grid
    record scatter as "hashgrid_scatter"; they are added into the phase
    (so a phase's time is the sum of its kernels' spans on both streams)."""
    out = {}
    for k, (ms, n) in phases.items():
        base = k[:-2] if k.endswith("/2") else k
        base = "hashgrid_backward" if base == "hashgrid_scatter" else base
        d = out.setdefault(base, dict(ms=0.0, launches=0))
        d["ms"] += ms / max(n, 1)
        d["launches"] = max(d["launches"], n)
    return out


def roofline(res, n_params, dtype, n_net=9408):
    """Achieved rate of every phase from its algorithmic bytes/FLOPs; the
    roofline object is the phase with the largest share of the step. With
    Adam fused into the hash-grid backward (the run's tuning pick), that
    phase's per-parameter bytes count the table's n_params - n_net
    parameters and Adam's the network's n_net."""
    S, R = res["samples_per_step"], res["rays_per_step"]
    fused = res.get("autotune", {}).get("picks", {}).get("hash_adam_fused", 0) == 1
    table = {}
    for ph, v in res["phases"].items():
        if ph not in ALGO:
            continue
        bound, ps, pr, pp = ALGO_FP32.get(ph, ALGO[ph]) if dtype == "fp32" else ALGO[ph]
        n_p = n_params
        if fused and ph in ALGO_FUSED:
            bound, ps, pr, pp = ALGO_FUSED[ph]
            n_p = n_params - n_net if ph == "hashgrid_backward" else n_net
            if ph == "hashgrid_backward" and dtype == "fp32":
                pp = 32
        units = ps * S + pr * R + pp * n_p
        sec = v["ms"] * 1e-3
        if bound == "hbm":
            ach, peak, unit = units / sec / 1e9, PEAK_HBM_GBS, "GB/s"
        else:
            ach, peak, unit = units / sec / 1e12, PEAK_TFLOPS[dtype], "TFLOP/s"
        table[ph] = dict(ms=round(v["ms"], 4), bound=bound, achieved=round(ach, 2), unit=unit,
                         frac=round(ach / peak, 4), algorithmic=units)
    dom = max((k for k in table if k in NETWORK_PHASES), key=lambda k: table[k]["ms"])
    d = table[dom]
    peak = PEAK_HBM_GBS if d["bound"] == "hbm" else PEAK_TFLOPS[dtype]
    kernels = KERNEL.get(dom, (dom,))
    traffic = measured_traffic(kernels, S)
    return dict(kernel="+".join(kernels), phase=dom, bound=d["bound"], achieved=d["achieved"], peak=peak,
                unit=d["unit"], frac=d["frac"], traffic=traffic, avg_launch_ms=d["ms"],
                algorithmic_per_launch=d["algorithmic"]), table


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "train_step_blob8_800x800_b262144.md")


def measured_traffic(kernels, samples_per_step):
    """HBM bytes per launch of the phase's kernels from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (its "HBM B/sample" column),
    scaled to this run's samples per launch; None when a kernel has no
    measurement."""
    if not kernels or not os.path.exists(TRAFFIC_PROFILE):
        return None
    per = {}
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0] in kernels:
                try:
                    per[cells[0]] = float(cells[3])
                except ValueError:
                    pass
    if any(k not in per for k in kernels):
        return None
    return int(sum(per.values()) * samples_per_step)


def render_rate(tn, synthetic, scene, trainer_tuple, n_frames=5):
    """Mrays/s of Renderer.submit on one 800x800 frame of the trained NeRF."""
    blobs, cams, images = scene
    mgr, nerf, tr = trainer_tuple
    cam = synthetic.to_pyturbonerf_cameras(tn, cams)[0]
    buf = tn.CPURenderBuffer()
    r = tn.Renderer()
    req = lambda: tn.RenderRequest(cam, [tn.Renderable(nerf)], buf)
    r.submit(req())  # warm-up
    t0 = time.perf_counter()
    for _ in range(n_frames):
        r.submit(req())
    dt = (time.perf_counter() - t0) / n_frames
    return RES * RES / dt / 1e6, 1e3 * dt


def timed_steps(tr, steps, graph=True):
    """(samples/s, ms/step) of `steps` training steps without host
    synchronisation inside, each the replay of a captured step by default."""
    tr.train_steps(1, read_metrics=False, graph=graph)  # captured before the clock starts
    tr.synchronize()
    s0 = tr.samples_trained()
    t0 = time.perf_counter()
    tr.train_steps(steps, read_metrics=False, graph=graph)
    tr.synchronize()
    wall = time.perf_counter() - t0
    return (tr.samples_trained() - s0) / wall, 1e3 * wall / steps


def distortion_rate(tr, steps, lam=1e-3):
    """The trained mixed-precision NeRF with the mip-NeRF 360 distortion loss
    switched on at weight `lam` (its forward and backward run every step)."""
    lam0 = tr.distortion_lambda
    tr.use_distortion_loss = True
    tr.distortion_lambda = lam
    tr.train_steps(16, read_metrics=False, graph=True)
    sps, ms = timed_steps(tr, steps)
    m = tr.train_step()
    tr.use_distortion_loss = False
    tr.distortion_lambda = lam0
    log(f"[bench distortion] lambda {lam}: {sps / 1e6:.2f} M samples/s, {ms:.3f} ms/step (loss with distortion "
        f"{m.loss:.5f})")
    return dict(samples_per_s=round(sps, 1), ms_per_step=round(ms, 4), distortion_lambda=lam)


# eight scenes: hash tables of 2^15 ... 2^21, 2^21 entries per level, blob
# fields of 4 to 32 blobs, uneven sample budgets summing to 2^19
MULTI_LOG2_T = [15, 16, 17, 18, 19, 20, 21, 21]
MULTI_BLOBS = [4, 8, 12, 16, 20, 24, 28, 32]
MULTI_BATCH = [8192, 16384, 32768, 65536, 131072, 131072, 65536, 73728]


def multi_scene(tn, synthetic, steps):
    """Eight NeRFs trained side by side: every round is one training step of
    each as replays of captured HIP graphs, each NeRF's own on its own
    streams (MultiTrainer). Aggregate samples/s after 128 warm-up steps each;
    the same rounds as one graph launch per round (MultiTrainer with
    one_launch) and as eight eager steps one after another are timed for
    comparison."""
    assert sum(MULTI_BATCH) == 1 << 19
    mgr = tn.NeRFManager()
    trainers = []
    for k in range(8):
        blobs, cams, images = synthetic.blob_dataset(seed=100 + k, n_cams=16, width=128, height=128,
                                                     n_blobs=MULTI_BLOBS[k])
        nerf = mgr.create()
        nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
        cfg = tn.TrainingConfig()
        cfg.network.precision = 1
        cfg.network.log2_hashmap_size = MULTI_LOG2_T[k]
        cfg.network.beta2 = 0.99
        cfg.rng_seed = 7 + k
        tr = tn.Trainer(nerf, cfg)
        tr.setup_data(batch_size=MULTI_BATCH[k])
        tr.load_images()
        trainers.append(tr)
    mt = tn.MultiTrainer(trainers)
    for i in range(8):
        mt.train_steps(16)
        for tr in trainers:
            tr.update_occupancy_grid(16 * (i + 1))
    tune = mt.autotune(rounds_per_trial=8, repetitions=3)
    mt.train_steps(2)  # both graphs captured before the clock starts
    rays = [tr.train_steps(0, read_metrics=True).n_rays for tr in trainers]

    one = tn.MultiTrainer(trainers, one_launch=True)
    one.train_steps(2)  # both round graphs captured before the clock starts

    def timed(runner):
        mt.synchronize()
        s0 = sum(tr.samples_trained() for tr in trainers)
        t0 = time.perf_counter()
        if runner is not None:
            runner.train_steps(steps)
        else:
            for _ in range(steps):
                for tr in trainers:
                    tr.train_steps(1, read_metrics=False)
        mt.synchronize()
        wall = time.perf_counter() - t0
        return sum(tr.samples_trained() for tr in trainers) - s0, wall

    samples, wall = timed(mt)
    samples_o, wall_o = timed(one)
    samples_e, wall_e = timed(None)
    captures = mt._graph_captures
    del mt, one
    for tr in trainers:
        tr.teardown()
    log(f"[bench multi-scene] {samples / wall / 1e6:.2f} M samples/s over 8 scenes ({1e3 * wall / steps:.3f} ms per "
        f"round, each NeRF's step graphs; one graph launch per round: {samples_o / wall_o / 1e6:.2f}; eight eager "
        f"steps: {samples_e / wall_e / 1e6:.2f} M samples/s), {samples / steps:.0f} samples per round, rays per "
        f"scene {rays}, {captures} captures, autotuned {tune['picks']}")
    return dict(samples_per_s=round(samples / wall, 1), ms_per_round=round(1e3 * wall / steps, 4),
                samples_per_round=round(samples / steps, 1), one_launch_samples_per_s=round(samples_o / wall_o, 1),
                eager_samples_per_s=round(samples_e / wall_e, 1), rays_per_scene=rays,
                log2_hashmap_size=MULTI_LOG2_T, batch=MULTI_BATCH, autotune=tune["picks"])


def large_scene(tn, synthetic):
    """Large-scene inference: a 16-unit box (5 cascades of 128^3) with 256
    small Gaussian blobs. The occupancy update of the random-init network
    (5 x 2^21 jittered points, density inference, the random-subset rule)
    is timed; the grid is then set from the blob field at cell centres
    (threshold 0.01) and a 1920x1080 frame of the network is rendered from
    a camera 24 units out."""
    import numpy as np
    import torch

    from turbonerf_amd import ops, scenes

    aabb, levels, seed = 16.0, 5, 404
    blobs = scenes.blob_field(seed, 256, aabb, sigma_range=(0.004, 0.012))
    bits = torch.zeros(128 ** 3, dtype=torch.uint8, device="cuda")
    dens = torch.zeros(levels * 128 ** 3, dtype=torch.float32, device="cuda")
    ops.blob_density_grid(ops.grid_struct(bits, dens, levels), torch.from_numpy(blobs.ravel()).cuda())
    dens_h = dens.cpu().numpy()
    cams = scenes.sphere_cameras(seed, 4, 24.0, 64, 64, 50.0, near=0.05, far=100.0)
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    ds = tn.Dataset(tn.BoundingBox(aabb), synthetic.to_pyturbonerf_cameras(tn, cams),
                    np.zeros((4, 64, 64, 4), np.uint8))
    nerf.attach_dataset(ds)
    cfg = tn.TrainingConfig()
    cfg.network.precision = 1
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=BATCH)
    tr.load_images()
    tr.train_step()  # a NeRF renders once it has trained a step
    tr.update_occupancy_grid(256)  # warm-up (allocations)
    tr.synchronize()
    t0 = time.perf_counter()
    n_upd = 4
    for k in range(n_upd):
        tr.update_occupancy_grid(272 + 16 * k)
    tr.synchronize()
    upd_ms = 1e3 * (time.perf_counter() - t0) / n_upd
    tr.set_occupancy_from_density(dens_h, 0.01)
    occ = float((tr.get_occupancy_bitfield() != 0).mean())
    W, H = 1920, 1080
    cam = synthetic.to_pyturbonerf_cameras(tn, scenes.sphere_cameras(seed + 1, 1, 24.0, W, H, 50.0))[0]
    buf = tn.CPURenderBuffer()
    r = tn.Renderer()
    req = lambda: tn.RenderRequest(cam, [tn.Renderable(nerf)], buf)
    r.submit(req())  # warm-up
    n_frames = 3
    t0 = time.perf_counter()
    for _ in range(n_frames):
        r.submit(req())
    dt = (time.perf_counter() - t0) / n_frames
    samples = int(r.last_sample_count)
    alpha = float(np.asarray(buf.get_rgba())[..., 3].mean())
    tr.teardown()
    out = dict(occupancy_update_ms=round(upd_ms, 3), occupied_cells=round(occ, 4),
               render_mrays_per_s=round(W * H / dt / 1e6, 2), render_ms_per_frame=round(1e3 * dt, 2),
               render_samples_per_frame=samples, render_msamples_per_s=round(samples / dt / 1e6, 1),
               mean_alpha=round(alpha, 4))
    log(f"[bench large scene] occupancy update {upd_ms:.2f} ms (5 cascades x 2^21 points); render "
        f"{W * H / dt / 1e6:.1f} Mrays/s ({1e3 * dt:.1f} ms per {W}x{H} frame, {samples} samples, "
        f"{occ:.3f} of cells occupied, mean alpha {alpha:.3f})")
    return out


def cpu_baseline(scene, bitfield, rays_per_step_gpu, args):
    """The oracle (CPU restatement of the same step, OpenMP over rays and
    samples, fp32) on a bounded sample: one training step of
    `args.cpu_samples` samples, 1 warm-up + median of `args.cpu_runs`."""
    import ctypes as C

    import numpy as np

    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle

    from turbonerf_amd import scenes

    blobs, cams, images = scene
    lib = oracle.lib()
    S = args.cpu_samples
    R = int(math.ceil(1.1 * rays_per_step_gpu * S / BATCH)) + 64
    g = __import__("turbonerf_amd._capi", fromlist=["HashGrid"]).HashGrid()
    b = math.exp((math.log(2048.0) - math.log(16.0)) / 15.0)
    lib.oc_hashgrid_layout(16, 2, 19, 16, b, C.addressof(g))
    rng = np.random.default_rng(0)
    table = rng.uniform(-1e-4, 1e-4, g.n_params).astype(np.float32)
    net = np.concatenate([rng.uniform(-a, a, n).astype(np.float32) for a, n in
                          ((math.sqrt(6 / 96), 2048), (math.sqrt(6 / 80), 1024), (math.sqrt(6 / 96), 2048),
                           (math.sqrt(6 / 128), 4096), (math.sqrt(6 / 67), 192))])
    params = np.concatenate([table, net])
    m1 = np.zeros_like(params)
    m2 = np.zeros_like(params)
    steps = np.zeros(params.size, np.uint32)
    grads = np.zeros_like(params)
    bbox = scenes.bbox(1.0)
    dt = float(np.float32(math.sqrt(3) / 1024))
    bits = np.ascontiguousarray(bitfield, np.uint8)
    offset = [0]

    def step():
        rnd = oracle.uniform(4 * R, 1234, offset[0])
        offset[0] += 4 * R
        rays = oracle.init_training_rays(R, R, N_CAMS, RES, RES, R / N_CAMS, bbox, cams, images, rnd)
        n_steps, alive = oracle.march_count(R, R, bbox, bits, 1, 128, 1 / 256, dt, dt, rays)
        offs, nr, ns, rc = oracle.scan_select(n_steps, R, S)
        if rc != 0 or ns == 0:
            return 0
        sm = oracle.march_generate(nr, R, ns, bbox, 1.0, bits, 1, 128, dt, dt, 1 / 256, rays, offs, alive, n_steps)
        feat = np.zeros(32 * ns, np.float32)
        lib.oc_hashgrid_fwd_f32(C.addressof(g), ns, ns, sm["pos"], params[:g.n_params], feat)
        feat = feat.reshape(32, ns)
        dir01 = sm["dir"].reshape(3, ns)
        dens, rgb, _, _ = oracle.mlp_train(params[g.n_params:], feat, dir01, None, None, None, 1.0, fp64=False)
        bg = rnd[R:4 * R]
        loss, dLd, dLc = oracle.composite_train(nr, R, ns, n_steps, offs, bg, rays["pix_rgba"], dens, rgb.ravel(),
                                                sm["dt"], fp64=False)
        gnet = np.zeros(9408, np.float32)
        _, _, dfeat, _ = oracle.mlp_train(params[g.n_params:], feat, dir01, dLc.reshape(3, ns), dLd, sm["t"], 1.0,
                                          fp64=False, dW=gnet)
        grads[:] = 0
        grads[g.n_params:] = gnet
        lib.oc_hashgrid_bwd_f32(C.addressof(g), ns, ns, sm["pos"], np.ascontiguousarray(dfeat.ravel()),
                                grads[:g.n_params])
        lr = 1e-2
        wl = params.copy()
        lib.oc_adam_step(params.size, g.n_params, params.size, 1.0, lr, 0.9, 0.99, 1e-15, 1e-6, params, wl, grads,
                         m1, m2, steps)
        return ns

    step()  # warm-up
    times, counts = [], []
    for _ in range(args.cpu_runs):
        t0 = time.perf_counter()
        n = step()
        times.append(time.perf_counter() - t0)
        counts.append(n)
    k = int(np.argsort(times)[len(times) // 2])
    rate = counts[k] / times[k]
    cores = oracle.num_threads()
    log(f"[bench cpu] {rate / 1e6:.3f} M samples/s on {cores} threads ({counts[k]} samples in {times[k]:.2f} s)")
    return dict(value=round(rate, 1), unit="samples/s", cores=cores, kind="port",
                sample=f"one training step of {counts[k]} samples ({R} rays) on the same scene and occupancy "
                       f"grid, fp32, median of {args.cpu_runs} after 1 warm-up")


def profile_steps(tn, synthetic, scene, args):
    """Training steps only (after the target renders), for rocprofv3: every
    dispatch after the last write_rgba_kernel belongs to a training step."""
    import torch

    mgr, nerf, tr = make_trainer(tn, synthetic, scene, 1)
    tr.train_steps(args.warmup + args.profile_steps, read_metrics=False)
    tr.synchronize()
    torch.cuda.synchronize()
    log(f"[bench profile] {args.warmup + args.profile_steps} training steps, {tr.samples_trained()} samples")
    tr.teardown()


def profile_render(tn, synthetic, scene, args):
    """Renders only (after training), for rocprofv3: every dispatch after the
    last adam_kernel belongs to one of the --profile-render frames."""
    import torch

    mgr, nerf, tr = make_trainer(tn, synthetic, scene, 1)
    step = 0
    while step < args.warmup:
        k = min(16, args.warmup - step)
        tr.train_steps(k, read_metrics=False)
        step += k
        if step % 16 == 0:
            tr.update_occupancy_grid(step)
    tr.train_step()
    tr.synchronize()
    torch.cuda.synchronize()
    mr, ms = render_rate(tn, synthetic, scene, (mgr, nerf, tr), n_frames=args.profile_render - 1)
    log(f"[bench profile] {args.profile_render} renders of {RES}x{RES}: {mr:.1f} Mrays/s ({ms:.2f} ms per frame)")
    tr.teardown()


def _short(name):
    base = name.split("(")[0]
    keys = {k for ks in KERNEL.values() for k in ks} | {
        "scan_tile_sums_kernel", "scan_tile_prefix_kernel", "scan_write_select_kernel", "uniform_kernel",
        "init_training_rays_kernel", "estimate_rays_kernel", "reduce_partials_kernel",
        "write_rgba_kernel", "cast_kernel", "composite_forward_kernel", "render_rays_kernel"}
    for key in sorted(keys, key=len, reverse=True):
        if key in base:
            return key
    return base[:48]


def summarize(paths, last=0, sps=0.0, after="write_rgba_kernel", units=0):
    """Markdown summary of a profiling run from rocprofv3 CSVs: the
    dispatches after the last dispatch of `after` (the target renders'
    write_rgba_kernel for --profile-steps, adam_kernel for --profile-render),
    or, with `last`, after the start of the last `last` Adam steps' window.
    Totals are per Adam step, or per `units` (frames) when given."""
    import csv
    import statistics as st

    def training_rows(path, ts_key):
        rows = list(csv.DictReader(open(path)))
        marks = [int(r[ts_key]) for r in rows if after in r["Kernel_Name"]]
        start = max(marks) if marks else 0
        if last:
            adam = sorted(int(r["End_Timestamp"]) for r in rows if "adam_kernel" in r["Kernel_Name"])
            if len(adam) > last:
                start = max(start, adam[-last - 1])
        return [r for r in rows if int(r[ts_key]) > start]

    trace = training_rows(paths[0], "Start_Timestamp")
    per = {}
    for r in trace:
        per.setdefault(_short(r["Kernel_Name"]), []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    steps = units or len(per.get("adam_kernel", [])) or 1
    total = sum(sum(v) for v in per.values())
    unit = "frame" if units else "step"
    out = [f"| kernel | calls | avg us | total us per {unit} | share |", "|---|---:|---:|---:|---:|"]
    for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| {k} | {len(v)} | {st.mean(v) / 1e3:.2f} | {sum(v) / 1e3 / steps:.2f} | "
                   f"{100 * sum(v) / total:.1f}% |")
    print("\n".join(out))
    if len(paths) >= 5:
        def counter(path):
            vals = {}
            for r in training_rows(path, "Start_Timestamp"):
                vals.setdefault(_short(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
            return {k: st.mean(v) for k, v in vals.items()}

        def copy_counter(path):
            rows = [r for r in csv.DictReader(open(path)) if "elementwise" in r["Kernel_Name"] or
                    "copy" in r["Kernel_Name"].lower()]
            return max(float(r["Counter_Value"]) for r in rows)

        fetch, write = counter(paths[1]), counter(paths[2])
        cf, cw = copy_counter(paths[3]), copy_counter(paths[4])
        probe = 256 * (1 << 20)  # --copy-probe-mib 256: bytes read and written per launch
        kf, kw = probe / cf, probe / cw
        print(f"\nCounter scale from the 256 MiB copy probe: FETCH_SIZE x {kf:.1f} B, WRITE_SIZE x {kw:.1f} B "
              f"per unit\n")
        head = "| kernel | HBM read MB/launch | HBM write MB/launch |"
        print(head + (" HBM B/sample |" if sps else ""))
        print("|---|---:|---:|" + ("---:|" if sps else ""))
        for k in sorted(fetch, key=lambda k: -fetch[k]):
            rd, wr = fetch[k] * kf, write.get(k, 0.0) * kw
            row = f"| {k} | {rd / 1e6:.2f} | {wr / 1e6:.2f} |"
            print(row + (f" {(rd + wr) / sps:.0f} |" if sps else ""))


def summarize_counters(paths, last=0, kernels=("mlp_forward_kernel", "mlp_backward_kernel"), n_cu=256, n_xcd=8):
    """Markdown table of the mean per-dispatch value of every counter in the
    rocprofv3 counter CSVs paths[1:] for `kernels`, over the same window as
    `summarize` (the last `last` training steps), with their average duration
    from the kernel trace paths[0]. Derived, where the counters are present:
    executed MFMA FLOP/s (SQ_INSTS_VALU_MFMA_MOPS_F16/_F32 count 512 FLOPs
    each) and MFMA busy cycles per SIMD-cycle of the kernel's duration
    (SQ_VALU_MFMA_BUSY_CYCLES, summed over the SIMDs, over the GPU's busy
    cycles x 4 SIMDs x n_cu; GRBM_GUI_ACTIVE reads the busy cycles summed
    over the n_xcd XCDs)."""
    import csv
    import statistics as st

    def window(rows):
        adam = sorted(int(r["End_Timestamp"]) for r in rows if "adam_kernel" in r["Kernel_Name"])
        start = adam[-last - 1] if last and len(adam) > last else 0
        return [r for r in rows if int(r["Start_Timestamp"]) > start]

    dur = {}
    for r in window(list(csv.DictReader(open(paths[0])))):
        dur.setdefault(_short(r["Kernel_Name"]), []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    vals = {}
    for p in paths[1:]:
        for r in window(list(csv.DictReader(open(p)))):
            k = _short(r["Kernel_Name"])
            if k in kernels:
                vals.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    names = sorted({c for v in vals.values() for c in v})
    print("| kernel | avg us | " + " | ".join(names) + " | MFMA TFLOP/s executed | MFMA busy |")
    print("|---|---:|" + "---:|" * len(names) + "---:|---:|")
    for k in kernels:
        if k not in vals:
            continue
        c = {n: st.mean(v) for n, v in vals[k].items()}
        us = st.mean(dur[k]) / 1e3 if k in dur else float("nan")
        mops = c.get("SQ_INSTS_VALU_MFMA_MOPS_F16", 0.0) + c.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0)
        tflops = f"{mops * 512 / (us * 1e-6) / 1e12:.1f}" if mops and us == us else "-"
        busy = "-"
        if "SQ_VALU_MFMA_BUSY_CYCLES" in c and c.get("GRBM_GUI_ACTIVE"):
            busy = f"{100 * c['SQ_VALU_MFMA_BUSY_CYCLES'] / (c['GRBM_GUI_ACTIVE'] / n_xcd * 4 * n_cu):.1f}%"
        print(f"| {k} | {us:.2f} | " + " | ".join(f"{c[n]:.4g}" if n in c else "-" for n in names)
              + f" | {tflops} | {busy} |")


def main():
    args = parse()
    if args.summarize_counters:
        summarize_counters(args.summarize_counters, args.summarize_last)
        return
    if args.summarize:
        summarize(args.summarize, args.summarize_last, args.samples_per_step, args.summarize_after,
                  args.summarize_units)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this build has no multi-GPU "
              f"path; the per-step gradient all-reduce it would need is described in the design notes)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import ops, synthetic

    torch.cuda.init()
    scene = make_scene()
    if args.profile_steps:
        profile_steps(tn, synthetic, scene, args)
        return
    if args.profile_render:
        profile_render(tn, synthetic, scene, args)
        return
    precs = {"mixed": [1], "fp32": [0], "both": [1, 0]}[args.precision]
    defaults = tn.get_tuning()  # the launch shapes before any tuning
    results = {}
    for p in precs:
        results[p] = run_precision(tn, synthetic, scene, p, args)
    main_p = precs[0]
    main = results[main_p]
    dtype = "fp16" if main_p == 1 else "fp32"
    n_params = ops.hashgrid_layout().n_params + 9408
    roof, table = roofline(main, n_params, dtype)
    extra = dict(phases_ms_per_step=table, occupancy_update_ms=round(main["occupancy_update_ms"], 3),
                 occupied_fraction=round(main["occupied_fraction"], 4), loss=round(main["loss"], 5),
                 samples_per_step=round(main["samples_per_step"], 1), rays_per_step=round(main["rays_per_step"], 1),
                 eager_samples_per_s=round(main["eager_samples_per_s"], 1),
                 eager_ms_per_step=round(main["eager_ms_per_step"], 4), autotune=main["autotune"])
    if 0 in results and main_p != 0:
        f = results[0]
        extra["fp32_samples_per_s"] = round(f["samples_per_s"], 1)
        extra["fp32_ms_per_step"] = round(f["ms_per_step"], 4)
        extra["mixed_over_fp32"] = round(main["samples_per_s"] / f["samples_per_s"], 3)
        extra["fp32_roofline"] = roofline(f, n_params, "fp32")[0]
        extra["fp32_phases_ms_per_step"] = roofline(f, n_params, "fp32")[1]
        extra["fp32_autotune"] = f["autotune"]
    # the renders run with the default launch shapes: the training picks are
    # tuned for 2^18-sample steps (e.g. a smaller marching grid), not for
    # render passes of up to 2^22 samples over 640K-2M rays
    for k, v in defaults.items():
        tn.set_tuning(k, v)
    if not args.no_render:
        mr, ms = render_rate(tn, synthetic, scene, main["trainer"])
        extra["render_mrays_per_s"] = round(mr, 2)
        extra["render_ms_per_frame"] = round(ms, 3)
        log(f"[bench] render {mr:.1f} Mrays/s ({ms:.2f} ms per {RES}x{RES} frame)")
    if not args.no_render:
        extra["large_scene_aabb16_5cascades_1920x1080"] = large_scene(tn, synthetic)
    for k, v in main["autotune"]["picks"].items():  # the distortion leg trains the main NeRF
        tn.set_tuning(k, v)
    if not args.no_extra_configs:
        if main_p == 1:
            extra["distortion_loss_mixed"] = distortion_rate(main["trainer"][2], args.steps)
        extra["multi_scene_8"] = multi_scene(tn, synthetic, 50)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, main["bitfield"], main["rays_per_step"], args)
    line = {
        "metric": METRIC,
        "value": round(main["samples_per_s"], 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: 8-blob Gaussian field rendered on the GPU at 800x800 x 100 cameras; random-init network",
        "config": {"workload": "nerf-train-blob8-800x800x100-batch2^18",
                   "hash_grid": "16 levels x 2 features, 2^19 entries/level, N 16->2048",
                   "mlp": "density 32-64-16, colour 32-64-64-3", "batch_samples": BATCH,
                   "optimizer": "Adam lr 1e-2 betas (0.9, 0.99) eps 1e-15", "loss": "smooth-L1",
                   "precision": "fp16 MFMA, fp32 accumulate, fp32 master weights" if main_p == 1 else "fp32"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    for p in results:
        results[p]["trainer"][2].teardown()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
