#!/usr/bin/env python3
"""Benchmark: tracked frames/s at 640x480 of learned-proposal + batched
Gauss-Newton tracking against the plain single-pair tracker, on the same
synthetic pairs, with convergence rates and pose errors.

Workload "learned" (default): 4096 pairs at 640x480 with large motions
(rotation <= 10 deg, translation <= 150 mm). A step is one pass of the hot path
over the batch, from raw 8-bit gray / 16-bit depth frames resident in HBM to
converged poses on the host: keyframe preparation (pyramids, gradient
histogram, point selection, Jacobians), the network's proposals (inputs,
forward pass, initial poses) and the coarse-to-fine photometric Gauss-Newton
loop of every pair in one persistent launch.
Workload "small": the same without the network, motions <= 3 deg / 50 mm,
identity start.

Plain = the single-pair tracker behind the PhotoTracker boundary (its own
residual and system kernels, the same per-point arithmetic, a different
summation order), one pair at a time from identity. A pair has converged when
its pose is within 0.1 deg / 1 mm of the FP64 oracle tracker started at the
true motion; convergence rates are measured on a sample of the pairs, and
`value` is the whole batch's frames/s times the sampled convergence rate
(converged frames/s). Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
# Algorithmic bytes per selected point per Gauss-Newton iteration: keyframe point
# xyz 12 + reference gray 4 + Jacobian 24 + four bilinear gray taps 16 + depth tap 4.
BYTES_PER_POINT_ITER = 60
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA
# Convolution FLOPs of the proposal network per pair (2 * Hout * Wout * Cout * 9 * Cin summed
# over the six layers at 80x60 input: 4-32-64-128-128-256-256)
CNN_FLOPS_PER_PAIR = 2 * 9 * (40 * 30 * 32 * 4 + 20 * 15 * 64 * 32 + 10 * 8 * 128 * 64 + 5 * 4 * 128 * 128
                              + 3 * 2 * 256 * 128 + 2 * 1 * 256 * 256)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(npairs, W, H, theta, tmax, seed0, nscenes=64):
    """Render npairs (reference, current) frames on the GPU into two device
    buffers each for gray and depth; returns buffers and ground-truth Trel."""
    from recon_amd import synth_lib
    from recon_amd.hip import DeviceBuffer, hip
    from recon_amd.synth import pinhole, sample_motion
    S = synth_lib()
    fx, fy, cx, cy = pinhole(W, H)
    npx = W * H
    g = DeviceBuffer(2 * npairs * npx)
    d = DeviceBuffer(2 * npairs * npx * 2)
    gts = []
    dp = ctypes.POINTER(ctypes.c_double)
    nscenes = min(npairs, nscenes)
    scenes = [S.rs_scene_create(ctypes.c_uint64(seed0 + k)) for k in range(nscenes)]
    for i in range(npairs):
        sc = scenes[i % nscenes]
        M = sample_motion(seed0 * 1000 + i, theta, tmax)
        T0 = np.eye(4).ravel().astype(np.float64)
        T1 = M.ravel().astype(np.float64)
        for f, T in ((0, T0), (1, T1)):
            k = 2 * i + f
            rc = S.rs_render_device(sc, T.ctypes.data_as(dp), W, H, fx, fy, cx, cy, ctypes.c_uint64(seed0 * 7919 + k),
                                    0.01, g.ptr + k * npx, d.ptr + k * npx * 2, None)
            if rc != 0:
                raise RuntimeError(f"render failed: {rc}")
        gts.append(np.linalg.inv(M))
    hip().hipDeviceSynchronize()
    for sc in scenes:
        S.rs_scene_destroy(sc)
    return g, d, np.array(gts)


def pose_errors(Trel_est, Trel_gt):
    from recon_amd.synth import pose_error
    e = np.array([pose_error(a, b) for a, b in zip(Trel_est, Trel_gt)])
    return e[:, 0], e[:, 1]


def measured_traffic(W, H, B, workload):
    """HBM bytes per launch of the tracking kernel at this workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE scaled by a
    copy kernel of known size); None where no summary exists for it."""
    path = os.path.join(ROOT, "profiles", f"kb_track_{W}x{H}_{B}pairs_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def copy_probe(mib):
    """One device-to-device copy kernel of a known size: reads and writes
    mib MiB each, so per-dispatch FETCH_SIZE / WRITE_SIZE of this kernel give
    the counters' scale in bytes."""
    import torch
    n = mib << 20
    x = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    x.fill_(7)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    assert int(y[n - 1]) == 7
    log(f"copy probe: {n} bytes copied")


def host_cpu():
    """The host's CPU model and the processing units this process may use
    (`nproc`, which honours the affinity mask and OMP_NUM_THREADS)."""
    import subprocess
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], stdout=subprocess.PIPE, text=True, check=True).stdout)
    except (OSError, ValueError, subprocess.CalledProcessError):
        nproc = len(os.sched_getaffinity(0))
    return model, nproc


def cpu_baseline(g, d, npairs, W, H, L, N, iters, cal, seconds, threads):
    """The oracle's FP32 tracker over a sample of the same pairs, one pair per
    thread at a time, for about `seconds` of wall time; frames/s over all
    threads."""
    import threading

    import oracle
    npx = W * H
    gh = g.to_array(np.uint8, 2 * npairs * npx).reshape(2 * npairs, H, W)
    dh = d.to_array(np.uint16, 2 * npairs * npx).reshape(2 * npairs, H, W)
    model, nproc = host_cpu()
    threads = max(1, min(threads, nproc)) if threads > 0 else nproc
    done = [0] * threads
    t0 = time.perf_counter()
    stop = t0 + seconds

    def work(k):
        i = k
        while time.perf_counter() < stop:
            p = i % npairs
            o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=L, iters=iters, nselect=N, dtype=np.float32)
            o.set_reference(gh[2 * p], dh[2 * p])
            o.set_current(gh[2 * p + 1], dh[2 * p + 1])
            o.track()
            done[k] += 1
            i += threads

    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    n = sum(done)
    return {"value": round(n / el, 2), "unit": "frames/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": nproc, "logical_cpus": os.cpu_count(),
            "sample": f"{n} pairs cycled from the first {npairs} {W}x{H} pairs of the batch, FP32 oracle tracker "
                      f"(the reference's kernels restated on the CPU), one pair per thread, {threads} threads, "
                      f"{el:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("learned", "small", "single", "sequence", "mixed", "pipeline"), default="learned",
                    help="learned: 4096 pairs at 640x480, large motions, network proposals + batched tracking "
                         "(the headline); small: the same batch from identity at small motions; single: one pair "
                         "at 320x240 on the single-pair tracker; sequence: 300 frames at 640x480, bi-objective, "
                         "keyframes every 10 frames fused with the next 5; mixed: 256 pairs of mixed sizes in one "
                         "persistent launch; pipeline: 2000 frames with learned warm starts and relocalisation")
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--plain-pairs", type=int, default=64)
    ap.add_argument("--width", type=int, default=640)
    ap.add_argument("--height", type=int, default=480)
    ap.add_argument("--theta", type=float, default=None, help="max rotation (deg) of a pair")
    ap.add_argument("--tmax", type=float, default=None, help="max translation (mm) of a pair")
    ap.add_argument("--weights", default=os.path.join(ROOT, "recon_amd", "data", "pose_cnn.npz"))
    ap.add_argument("--oracle-pairs", type=int, default=256,
                    help="pairs checked against the FP64 oracle tracker started at the true motion")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="wall time of the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU threads (0: nproc)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--inflight", type=int, default=1, choices=(1, 2),
                    help="learned workload: batches in flight (2: the next step's preparation overlaps this step's tracking)")
    ap.add_argument("--extras", action="store_true",
                    help="also run the single-pair and sequence cases of the earlier build's extra leg "
                         "(--workload single / sequence give them their own lines)")
    ap.add_argument("--track-only", action="store_true",
                    help="prepare the batch and run --steps tracking launches only (for counter collection)")
    ap.add_argument("--tsdf", action="store_true",
                    help="time TSDF fusion and ray casting (256^3 voxels, 320x240) instead of tracking")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run one device copy of MIB MiB and exit (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent: shard them across GPUs and gather poses)", file=sys.stderr)
        sys.exit(2)

    if args.tsdf:
        tsdf_bench(args)
        return
    if args.workload == "single":
        single_bench(args)
        return
    if args.workload == "sequence":
        sequence_bench(args)
        return
    if args.workload == "mixed":
        mixed_bench(args)
        return
    if args.workload == "pipeline":
        pipeline_bench(args)
        return

    from recon_amd import PHOTOMETRIC
    from recon_amd.hip import Event, stream_sync
    from recon_amd.synth import calib_block
    from recon_amd.tracker import BatchTracker, PhotoTracker

    learned = args.workload == "learned"
    theta = args.theta if args.theta is not None else (10.0 if learned else 3.0)
    tmax = args.tmax if args.tmax is not None else (150.0 if learned else 50.0)
    W, H, B = args.width, args.height, args.pairs
    iters = (1, 3, 10, 5)
    N = 8192
    L = 4 if W == 640 else 3
    t0 = time.time()
    g, d, gt = make_inputs(B, W, H, theta, tmax, args.seed, nscenes=B if learned else 64)
    log(f"inputs: {B} pairs {W}x{H} (<= {theta} deg, <= {tmax} mm) rendered in {time.time() - t0:.1f}s")
    npx = W * H
    cal = calib_block(W, H)
    pairs = [dict(ref_gray=g.ptr + 2 * i * npx, ref_depth=d.ptr + 2 * i * npx * 2, cur_gray=g.ptr + (2 * i + 1) * npx,
                  cur_depth=d.ptr + (2 * i + 1) * npx * 2, width=W, height=H, calib=cal) for i in range(B)]
    bt = BatchTracker(B, W, H, mode=PHOTOMETRIC, n_select=N, iters=iters)
    stream = bt.L.rt_batch_stream(bt.h)
    prepared = bt.prepare(pairs)
    net = None
    if learned:
        from recon_amd.cnn import PoseNet
        net = PoseNet.load(args.weights, max_batch=B)

    def step():
        if learned:
            bt.prepare_only(prepared)
            bt.propose(net)
            return bt.retrack(None, B)
        return bt.track(prepared)

    if args.track_only:
        step()
        for _ in range(args.steps):
            bt.L.rt_batch_retrack(bt.h, None, None)
        stream_sync(stream)
        log(f"track-only: {args.steps} tracking launches over {B} pairs done")
        return

    # ---- plain single-pair tracker (its own kernels, rt_kernels.hip; one pair at a time, identity start)
    P = min(args.plain_pairs, B)
    tr = PhotoTracker(W, H, cal, mode=PHOTOMETRIC, n_select=N, iters=iters)
    plain_T = []

    def plain_pair(i):
        tr.set_reference_points(pairs[i]["ref_gray"], pairs[i]["ref_depth"], mode=PHOTOMETRIC, force=True, on_device=True)
        tr.set_current_image(pairs[i]["cur_gray"], pairs[i]["cur_depth"], on_device=True)
        tr.track()
        return tr.trel().astype(np.float64)

    for i in range(min(4, P)):
        plain_pair(i)
    tr.synchronize()
    t1 = time.perf_counter()
    for i in range(P):
        plain_T.append(plain_pair(i))
    tr.synchronize()
    plain_s = time.perf_counter() - t1
    plain_fps = P / plain_s
    log(f"plain single-pair: {P} pairs in {plain_s * 1e3:.1f} ms -> {plain_fps:.1f} frames/s")

    # ---- batched identity start (untimed, for the convergence comparison)
    bt.prepare_only(prepared)
    ident, _, _, _ = bt.retrack(np.tile(np.eye(4).ravel(), B), B)

    # ---- the timed step
    for _ in range(max(1, args.warmup)):
        Trel_b, _, rms, st = step()
    stream_sync(stream)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        Trel_b, _, rms, st = step()
    stream_sync(stream)
    step_s = (time.perf_counter() - t2) / args.steps
    serial_step_s = step_s
    frames_per_s = B / step_s

    # ---- dominant kernel (the persistent tracking launch), HIP events on its stream
    ev0, ev1, ev2 = Event(), Event(), Event()
    reps = 3
    ev0.record(stream)
    for _ in range(reps):
        bt.L.rt_batch_retrack(bt.h, None, None)
    ev1.record(stream)
    ev1.synchronize()
    k_ms = ev0.elapsed_ms(ev1) / reps
    alg_bytes = B * sum(iters[:L]) * N * BYTES_PER_POINT_ITER
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    cnn = None
    if learned:
        ev0.record(stream)
        for _ in range(reps):
            bt.propose(net)
        ev2.record(stream)
        ev2.synchronize()
        c_ms = ev0.elapsed_ms(ev2) / reps
        # the convolutions alone (forward + head on the network's inputs), on
        # the network's own stream with nothing else running
        from recon_amd.cnn import pair_array
        from recon_amd.hip import DeviceBuffer
        xin, yout = DeviceBuffer(2 * net.input_elems * B), DeviceBuffer(24 * B)
        net.make_inputs(pair_array(pairs), B, xin.ptr)
        nstream = net.L.rt_cnn_stream(net.h)
        net.forward(B, xin.ptr, yout.ptr)
        ev0.record(nstream)
        for _ in range(reps):
            net.forward(B, xin.ptr, yout.ptr)
        ev2.record(nstream)
        ev2.synchronize()
        f_ms = ev0.elapsed_ms(ev2) / reps
        flops = B * CNN_FLOPS_PER_PAIR
        cnn = {"bound": "mfma", "achieved": round(flops / (f_ms * 1e-3) / 1e12, 2), "peak": BF16_PEAK_TFLOPS,
               "unit": "TFLOP/s", "frac": round(flops / (f_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 5),
               "ms": round(f_ms, 3), "flops_per_pair": CNN_FLOPS_PER_PAIR, "proposal_ms": round(c_ms, 3),
               "note": "ms: the six convolutions + head of the batch (forward on the network's stream); "
                       "proposal_ms: the whole proposal step on the batch (forward + poses, on the inputs "
                       "the batch preparation built; the network's own input builder, used before the "
                       "first proposal of a batch, streams both frames, 7.5 GB)"}
        del xin, yout

    # ---- two batches in flight (learned): the host's preparation of step k+1
    # and its launches overlap step k's tracking; every step still prepares,
    # proposes and tracks its 4096 pairs and returns their poses to the host
    inflight = None
    if learned and args.inflight == 2:
        from recon_amd.cnn import PoseNet
        bt2 = BatchTracker(B, W, H, mode=PHOTOMETRIC, n_select=N, iters=iters)
        prepared2 = bt2.prepare(pairs)
        bt2.track(prepared2)  # the second batch's variant is tuned with the GPU otherwise idle
        net2 = PoseNet.load(args.weights, max_batch=B)
        lanes = [(bt, prepared, net), (bt2, prepared2, net2)]

        def launch(k):
            b, pr, n = lanes[k % 2]
            b.prepare_only(pr)
            b.propose(n)
            b.retrack_async()

        launch(0)
        launch(1)
        lanes[0][0].results(B)
        lanes[1][0].results(B)
        t4 = time.perf_counter()
        launch(0)
        for k in range(1, args.steps):
            launch(k)
            Trel_p, _, _, _ = lanes[(k - 1) % 2][0].results(B)
        Trel_p, _, _, _ = lanes[(args.steps - 1) % 2][0].results(B)
        pipe_s = (time.perf_counter() - t4) / args.steps
        same = bool(np.array_equal(Trel_p, Trel_b))
        inflight = {"ms_per_step": round(pipe_s * 1e3, 3), "serial_ms_per_step": round(serial_step_s * 1e3, 3),
                    "poses_identical_to_serial": same}
        log(f"two batches in flight: {pipe_s * 1e3:.2f} ms/step (serial {serial_step_s * 1e3:.2f}); "
            f"poses identical to the serial steps: {same}")
        step_s = pipe_s
        frames_per_s = B / step_s
        del lanes, bt2, net2

    # ---- convergence and errors against the FP64 oracle tracker started at the
    # true motion (the convergence reference), on the first S pairs of the batch
    S = max(P, min(args.oracle_pairs, B))
    t3 = time.perf_counter()
    ref64 = fp64_reference(g, d, gt, S, W, H, L, N, iters, cal, args.cpu_threads)
    log(f"FP64 oracle from the true motion: {S} pairs in {time.perf_counter() - t3:.1f} s")

    def conv(T, ref):
        r, t = pose_errors(T, ref)
        ok = (r <= 0.1) & (t <= 1.0)
        return ok, r, t
    ok_b, rb, tb = conv(Trel_b[:S], ref64)
    ok_i, _, _ = conv(ident[:S], ref64)
    ok_p, rp, tp = conv(np.array(plain_T), ref64[:P])
    rate = float(ok_b.mean())
    value = frames_per_s * rate
    # the batched and the plain tracker sum the normal equations in different
    # orders, so they agree to rounding where both converge; a diverging pair
    # amplifies the difference, so the comparison is over the pairs both converge on
    both = ok_i[:P] & ok_p
    ra, ta = pose_errors(ident[:P][both], np.array(plain_T)[both]) if both.any() else (np.zeros(1), np.zeros(1))
    rtb, ttb = pose_errors(Trel_b[:S], gt[:S])
    plain_conv_fps = plain_fps * float(ok_p.mean())
    log(f"{'learned+batched' if learned else 'batched'}: {B} pairs, {step_s * 1e3:.2f} ms/step -> {frames_per_s:.0f} "
        f"frames/s, {value:.0f} converged frames/s ({value / plain_fps:.1f}x plain); kb_track {k_ms:.2f} ms, "
        f"{achieved:.0f} GB/s algorithmic")
    log(f"converged (<=0.1 deg / 1 mm of the FP64 oracle from the truth, {S} pairs): "
        f"{'learned+batched' if learned else 'batched'} {rate * 100:.1f}%, "
        f"batched from identity {ok_i.mean() * 100:.1f}%, plain from identity {ok_p.mean() * 100:.1f}% (of {P}); "
        f"batched vs plain from identity, {int(both.sum())} pairs both converge on: max {ra.max():.2e} deg {ta.max():.2e} mm")
    fp64 = {"pairs": S, "converged": int(ok_b.sum()),
            "converged_max_deg": float(rb[ok_b].max()) if ok_b.any() else None,
            "converged_max_mm": float(tb[ok_b].max()) if ok_b.any() else None,
            "converged_median_deg": float(np.median(rb[ok_b])) if ok_b.any() else None,
            "converged_median_mm": float(np.median(tb[ok_b])) if ok_b.any() else None,
            # error of every sampled result against the FP64 oracle (deg, mm), pair order
            "per_pair_deg": [round(float(x), 5) for x in rb], "per_pair_mm": [round(float(x), 4) for x in tb]}
    log(f"vs FP64 oracle (converged pairs): max {fp64['converged_max_deg']} deg {fp64['converged_max_mm']} mm")

    # ---- CPU baseline: the oracle's FP32 transcription on the host cores
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(g, d, min(B, 64), W, H, L, N, iters, cal, args.cpu_seconds, args.cpu_threads)
        log(f"cpu baseline: {cpu['value']} frames/s ({cpu['sample']})")

    extras = None
    if args.extras:
        extras = {"single_pair_320x240": extra_single_pair(), "sequence_640x480": extra_sequence()}
        for k, v in extras.items():
            log(f"{k}: {v}")

    name = "learned proposal + batched Gauss-Newton" if learned else "batched Gauss-Newton from identity"
    out = {
        "metric": "tracked frames/s at 640x480 (learned+batched vs plain GN) at matched pose error",
        "value": round(value, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" + ("+bf16" if learned else ""),
        "data": "synthetic",
        "config": {"workload": f"{name}: {B} frame pairs at {W}x{H}, rotation <= {theta} deg, translation <= {tmax} mm, "
                               f"{N} points, {L} levels, iterations {list(iters[:L])}",
                   "pairs": B, "width": W, "height": H, "points": N, "levels": L, "theta_deg": theta, "tmax_mm": tmax},
        "frames_per_s_all_pairs": round(frames_per_s, 1),
        "batches_in_flight": inflight,
        "plain_fps": round(plain_fps, 1),
        "plain_converged_fps": round(plain_conv_fps, 1),
        # converged frames/s over the plain tracker's frames/s, whether or not its pairs converge
        "speedup_vs_plain": round(value / plain_fps, 2),
        "convergence": {"criterion": "within 0.1 deg and 1 mm of the FP64 oracle tracker started at the true motion",
                        "sample": S, "this": round(rate, 4), "batched_from_identity": round(float(ok_i.mean()), 4),
                        "plain_from_identity": round(float(ok_p.mean()), 4), "plain_sample": P},
        "pose_error": {"converged_median_deg_vs_truth": float(np.median(rtb[ok_b])) if ok_b.any() else None,
                       "converged_median_mm_vs_truth": float(np.median(ttb[ok_b])) if ok_b.any() else None,
                       "batched_vs_plain_identity_pairs": int(both.sum()),
                       "batched_vs_plain_identity_max_deg": float(ra.max()),
                       "batched_vs_plain_identity_max_mm": float(ta.max()), "vs_fp64_oracle": fp64},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(W, H, B, args.workload),
                     "algorithmic_bytes": alg_bytes, "kernel": "kb_track (persistent Gauss-Newton)",
                     "kernel_ms": round(k_ms, 3)},
        "cnn_roofline": cnn,
        "cpu_baseline": cpu,
        "other_workloads": extras,
    }
    print(json.dumps(out), flush=True)


def _line(metric, value, unit, args, steps, step_s, dtype, config, **extra):
    out = {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def single_bench(args):
    """One pair at 320x240 (fx = fy = 262.5, rotation <= 2 deg, translation <=
    30 mm) on the single-pair tracker: 3 levels x 10 Gauss-Newton iterations,
    256-bin MAD Tukey weights, 8192 points. A step = keyframe preparation +
    tracking of the pair from frames resident in HBM to the pose on the host.
    Roofline: 60 B per point-iteration plus 12.3 B per level-0 pixel of
    conversion and pyramids (15.7 MB per pair)."""
    import oracle
    from recon_amd import PHOTOMETRIC, synth
    from recon_amd.hip import DeviceBuffer, Event
    from recon_amd.tracker import PhotoTracker
    W, H, N, iters = 320, 240, 8192, (10, 10, 10, 0)
    g0, d0, g1, d1, Tgt = synth.make_pair(1, W, H, 2.0, 30.0)
    cal = synth.calib_block_np(W, H)
    bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
    tr = PhotoTracker(W, H, cal, mode=PHOTOMETRIC, n_select=N, iters=iters, mad_bins=256)
    s = tr.L.rt_stream(tr.h)

    def step():
        tr.set_reference_points(bufs[0].ptr, bufs[1].ptr, mode=PHOTOMETRIC, force=True, on_device=True)
        tr.set_current_image(bufs[2].ptr, bufs[3].ptr, on_device=True)
        tr.track()

    for _ in range(max(1, args.warmup) * 10):
        step()
    tr.synchronize()
    steps = max(args.steps, 1) * 50
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    tr.synchronize()
    step_s = (time.perf_counter() - t0) / steps
    # the step's device time on the tracker's stream (HIP events)
    e0, e1 = Event(), Event()
    e0.record(s)
    for _ in range(20):
        tr.set_reference_points(bufs[0].ptr, bufs[1].ptr, mode=PHOTOMETRIC, force=True, on_device=True)
        tr.set_current_image(bufs[2].ptr, bufs[3].ptr, on_device=True)
        tr.L.rt_track(tr.h)
    e1.record(s)
    e1.synchronize()
    dev_ms = e0.elapsed_ms(e1) / 20
    # host frames in (PCIe inclusive), for reference
    t1 = time.perf_counter()
    for _ in range(20):
        tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
        tr.set_current_image(g1, d1)
        tr.track()
    host_s = (time.perf_counter() - t1) / 20
    T = tr.trel().astype(np.float64)
    res = {}
    for dt in (np.float32, np.float64):
        o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=3, iters=iters, nselect=N, mad_bins=256, dtype=dt)
        o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
        res[dt] = o.poses()[0]
    e32 = synth.pose_error(T, res[np.float32])
    e64 = synth.pose_error(T, res[np.float64])
    et = synth.pose_error(T, Tgt)
    alg = sum(iters[:3]) * N * BYTES_PER_POINT_ITER + 12.3 * W * H
    achieved = alg / (dev_ms * 1e-3) / 1e9
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_sample(lambda: _oracle_pair(oracle, W, H, cal, g0, d0, g1, d1, iters, N), args.cpu_seconds,
                         args.cpu_threads, "the pair at 320x240, FP32 oracle tracker (the reference's kernels restated "
                                           "on the CPU), one pair per thread")
    log(f"single pair 320x240: {1 / step_s:.0f} pairs/s (device {dev_ms:.3f} ms per pair, host frames in "
        f"{1 / host_s:.0f} pairs/s); vs FP32 oracle {e32[0]:.2e} deg {e32[1]:.2e} mm, vs FP64 oracle {e64[0]:.2e} deg "
        f"{e64[1]:.2e} mm, vs truth {et[0]:.3f} deg {et[1]:.2f} mm")
    _line("tracked frames/s, one pair at 320x240 on the single-pair tracker", 1.0 / step_s, "frames/s", args, steps,
          step_s, "f32+f64",
          {"workload": "plain photometric tracking of one pair at 320x240 (fx = fy = 262.5), rotation <= 2 deg, "
                       "translation <= 30 mm, 3 levels x 10 iterations, 256-bin MAD Tukey, 8192 points; frames "
                       "resident in HBM, keyframe preparation in every step", "width": W, "height": H, "points": N},
          host_frames_fps=round(1.0 / host_s, 1),
          pose_error={"vs_fp32_oracle_deg": e32[0], "vs_fp32_oracle_mm": e32[1], "vs_fp64_oracle_deg": e64[0],
                      "vs_fp64_oracle_mm": e64[1], "vs_truth_deg": et[0], "vs_truth_mm": et[1]},
          roofline={"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "algorithmic_bytes": alg,
                    "kernel": "the whole step (keyframe preparation + 30 iterations), HIP events on the tracker's "
                              "stream", "kernel_ms": round(dev_ms, 4)},
          cpu_baseline=cpu)


def _oracle_pair(oracle, W, H, cal, g0, d0, g1, d1, iters, N):
    o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=3, iters=iters, nselect=N, mad_bins=256,
                       dtype=np.float32)
    o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
    return 1


def cpu_sample(work, seconds, threads, what):
    """Runs work() (returning the frames it tracked) on `threads` threads (0:
    nproc) for about `seconds`; frames/s over all threads."""
    import threading
    model, nproc = host_cpu()
    threads = max(1, min(threads, nproc)) if threads > 0 else nproc
    done = [0] * threads
    t0 = time.perf_counter()
    stop = t0 + seconds

    def run(k):
        while time.perf_counter() < stop:
            done[k] += work()

    ts = [threading.Thread(target=run, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    n = sum(done)
    return {"value": round(n / el, 3), "unit": "frames/s", "cores": threads, "kind": "port", "cpu_model": model,
            "nproc": nproc, "logical_cpus": os.cpu_count(), "sample": f"{n} frames: {what}; {threads} threads, "
                                                                     f"{el:.1f} s wall"}


def sequence_bench(args):
    """300 frames at 640x480 (fx = fy = 525) along a seeded smooth trajectory
    (<= 1.5 deg, <= 20 mm per frame), bi-objective (lambda 0.5) tracking with
    61440 photometric points, 4 levels, iterations {1,3,10,5}; a keyframe every
    10 frames, fused with the next 5. A step = the whole sequence, frames
    resident in HBM, poses to the host at the end (rt_track_sequence).
    Roofline: per frame and iteration 60 B per photometric point and 80 B per
    valid geometric point of the level (the last frame's counts), plus 12.3 B
    per pixel of conversion and pyramids and the keyframe preparation
    amortised over its 10 frames."""
    import oracle
    from recon_amd import BIOBJECTIVE, synth, synth_lib
    from recon_amd.hip import DeviceBuffer, Event, hip
    from recon_amd.pipeline import relative_pose_error
    from recon_amd.tracker import PhotoTracker
    W, H, n, seed, N, iters = 640, 480, 300, 5, 61440, (1, 3, 10, 5)
    S = synth_lib()
    poses = synth.trajectory(seed * 13 + 5, n, 1.5, 20.0, 0, 1.0)
    fx, fy, cx, cy = synth.pinhole(W, H)
    npx = W * H
    g, d = DeviceBuffer(n * npx), DeviceBuffer(n * npx * 2)
    sc = S.rs_scene_create(ctypes.c_uint64(seed))
    dp = ctypes.POINTER(ctypes.c_double)
    for k in range(n):
        T = np.ascontiguousarray(poses[k], np.float64).ravel()
        S.rs_render_device(sc, T.ctypes.data_as(dp), W, H, fx, fy, cx, cy, ctypes.c_uint64(seed * 1000 + k), 0.01,
                           g.ptr + k * npx, d.ptr + k * npx * 2, None)
    hip().hipDeviceSynchronize()
    S.rs_scene_destroy(sc)
    cal = synth.calib_block_np(W, H)
    tr = PhotoTracker(W, H, cal, mode=BIOBJECTIVE, n_select=N, lam=0.5, iters=iters)
    s = tr.L.rt_stream(tr.h)
    for _ in range(max(1, args.warmup)):
        tr.track_sequence(g.ptr, d.ptr, 30, 10, 5)
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(steps):
        P = tr.track_sequence(g.ptr, d.ptr, n, 10, 5)
    step_s = (time.perf_counter() - t0) / steps
    fps = n / step_s
    e0, e1 = Event(), Event()
    e0.record(s)
    tr.track_sequence(g.ptr, d.ptr, n, 10, 5)
    e1.record(s)
    e1.synchronize()
    dev_ms = e0.elapsed_ms(e1)
    ng = tr.debug_read("ICP_COUNT", 0, np.int32, 4)
    per_frame = sum(iters[3 - l] * (BYTES_PER_POINT_ITER * N + 80 * int(ng[l])) for l in range(4)) + 12.3 * W * H
    per_frame += (W * H * 8 + 4 * 40 * N) / 10.0
    achieved = per_frame * n / (dev_ms * 1e-3) / 1e9
    et = np.array([synth.pose_error(P[k], poses[k]) for k in range(n)])
    rpe_mm, rpe_deg = relative_pose_error(P, poses, 100)
    cpu = None
    if args.cpu_seconds > 0:
        gh = g.to_array(np.uint8, 31 * npx).reshape(31, H, W)
        dh = d.to_array(np.uint16, 31 * npx).reshape(31, H, W)

        def work():
            # frames 1..10 of the sequence against keyframe 0, fused as on the GPU
            o = oracle.Tracker(W, H, cal, mode=oracle.BIOBJECTIVE, levels=4, iters=iters, nselect=N, lam=0.5,
                               dtype=np.float32)
            o.set_reference(gh[0], dh[0])
            for k in range(1, 3):
                o.set_current(gh[k], dh[k]); o.track(); o.fuse_current()
            return 2
        cpu = cpu_sample(work, args.cpu_seconds, args.cpu_threads,
                         "frames 1-2 of the sequence against keyframe 0 with fusion, bi-objective, 61440 points, FP32 "
                         "oracle tracker (the reference's kernels restated on the CPU), one sequence per thread")
    log(f"sequence 640x480: {fps:.1f} frames/s ({dev_ms / n:.3f} ms per frame on the device); {achieved:.0f} GB/s "
        f"algorithmic; final error {et[-1, 0]:.3f} deg {et[-1, 1]:.2f} mm; RPE(100) {rpe_mm:.2f} mm {rpe_deg:.3f} deg; "
        f"ICP points per level {list(map(int, ng))}")
    _line("tracked frames/s of a 300-frame bi-objective sequence at 640x480", fps, "frames/s", args, steps, step_s,
          "f32+f64",
          {"workload": "bi-objective (lambda 0.5) keyframe tracking of a seeded 300-frame sequence at 640x480, "
                       "fx = fy = 525, <= 1.5 deg / 20 mm per frame, 61440 photometric points, 4 levels, iterations "
                       "[1, 3, 10, 5], keyframe every 10 frames fused with the next 5; frames resident in HBM",
           "frames": n, "width": W, "height": H, "points": N},
          drift={"final_deg": et[-1, 0], "final_mm": et[-1, 1], "rpe100_mm": rpe_mm, "rpe100_deg": rpe_deg,
                 "drift_mm_per_100_frames": et[-1, 1] / (n / 100.0), "max_deg": float(et[:, 0].max()),
                 "max_mm": float(et[:, 1].max())},
          roofline={"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "algorithmic_bytes_per_frame": per_frame,
                    "icp_points_per_level": [int(x) for x in ng],
                    "kernel": "the whole sequence (every launch of rt_track_sequence), HIP events on the tracker's "
                              "stream", "kernel_ms": round(dev_ms, 3)},
          cpu_baseline=cpu)


def tsdf_bench(args):
    """TSDF fusion and ray casting at the reference's grid (256^3 voxels,
    half-extent 1500 mm, 320x240 depth): one step = distance map from 16-bit
    depth + update + ray cast with refinement and shading, for one frame of a
    synthetic sequence whose depth frames are resident on the device. Kernel
    times come from HIP events on the grid's stream."""
    from recon_amd import lib
    from recon_amd.hip import DeviceBuffer, Event
    from recon_amd.synth import make_sequence, pinhole
    from recon_amd.tsdf import VoxelGrid, default_params
    W, H, n = 320, 240, 16
    # noise-free depth: neighbouring depths agree, so every voxel in front of the
    # surface inside the view is updated (the densest update of the sequence)
    _, depth, poses = make_sequence(11, n, W, H, 1.0, 15.0, noise=0.0)
    fx, fy, cx, cy = pinhole(W, H)
    K = np.array([-fx / W, 0, cx / W, 0, fy / H, cy / H, 0, 0, 1], np.float32)
    Kp = K.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    Ts = [np.ascontiguousarray(p, np.float32).ravel() for p in poses]
    Tp = [t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for t in Ts]
    frames = DeviceBuffer.from_array(depth)
    fptr = [frames.ptr + k * W * H * 2 for k in range(n)]
    P = default_params()
    g = VoxelGrid(P)
    L = lib()
    xyz, nrm, shade = DeviceBuffer(W * H * 16), DeviceBuffer(W * H * 16), DeviceBuffer(W * H * 16)
    # voxels one frame updates (fresh grid): 8 B read + 8 B written each
    upd = []
    for k in range(0, n, 4):
        g.reset()
        L.rt_tsdf_set_depth_u16(g.h, fptr[k])
        L.rt_tsdf_update(g.h, Tp[k], Kp)
        upd.append(int((g.read()[..., 1] > 0).sum()))
    upd_voxels = float(np.mean(upd))
    g.reset()
    s = g.stream

    def step(k):
        L.rt_tsdf_set_depth_u16(g.h, fptr[k % n])
        L.rt_tsdf_update(g.h, Tp[k % n], Kp)
        L.rt_tsdf_raycast(g.h, Tp[k % n], Kp, 1, xyz.ptr, nrm.ptr, shade.ptr)

    for k in range(max(1, args.warmup) * 4):
        step(k)
    g.synchronize()
    steps = max(args.steps, 1) * 10
    t0 = time.perf_counter()
    for k in range(steps):
        step(k)
    g.synchronize()
    step_s = (time.perf_counter() - t0) / steps
    # kernel times on the grid's stream
    e = [Event() for _ in range(4)]
    reps = 20
    tu = tr = 0.0
    for k in range(reps):
        L.rt_tsdf_set_depth_u16(g.h, fptr[k % n])
        e[0].record(s)
        L.rt_tsdf_update(g.h, Tp[k % n], Kp)
        e[1].record(s)
        L.rt_tsdf_raycast(g.h, Tp[k % n], Kp, 1, xyz.ptr, nrm.ptr, shade.ptr)
        e[2].record(s)
        e[2].synchronize()
        tu += e[0].elapsed_ms(e[1]) / reps
        tr += e[1].elapsed_ms(e[2]) / reps
    alg = upd_voxels * 16
    nvox = P.resolution ** 3
    out = {
        "metric": "TSDF frames/s (fuse + ray cast) at 320x240 into 256^3 voxels",
        "value": round(1.0 / step_s, 1), "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "TSDF update + ray cast (refined, shaded) per frame, 256^3 voxels, half-extent 1500 mm, "
                               "320x240 noise-free 16-bit depth resident on the device, 16-frame synthetic sequence",
                   "resolution": P.resolution, "width": W, "height": H},
        "update": {"ms": round(tu, 4), "voxels": nvox, "voxels_per_s": round(nvox / (tu * 1e-3), 1),
                   "updated_voxels": upd_voxels,
                   "roofline": {"bound": "hbm", "achieved": round(alg / (tu * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(alg / (tu * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                "algorithmic_bytes": alg,
                                "note": "8 B read + 8 B written per updated voxel; voxels left alone are not read"}},
        "raycast_ms": round(tr, 4),
    }
    print(json.dumps(out), flush=True)


def extra_single_pair():
    """One pair at 320x240 (fx = fy = 262.5), motion <= 2 deg / 30 mm: 3 levels x
    10 iterations, 256-bin MAD Tukey; GPU frames/s of the single-pair tracker,
    error against the FP64 oracle and the oracle's own time."""
    import oracle
    from recon_amd import PHOTOMETRIC, synth
    from recon_amd.tracker import PhotoTracker
    W, H = 320, 240
    g0, d0, g1, d1, Tgt = synth.make_pair(1, W, H, 2.0, 30.0)
    cal = synth.calib_block_np(W, H)
    kw = dict(mode=PHOTOMETRIC, n_select=8192, iters=(10, 10, 10, 0), mad_bins=256)
    tr = PhotoTracker(W, H, cal, **kw)
    tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
    tr.set_current_image(g1, d1)
    tr.track()
    reps = 50
    t0 = time.perf_counter()
    for _ in range(reps):
        tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
        tr.set_current_image(g1, d1)
        tr.track()
    tr.synchronize()
    fps = reps / (time.perf_counter() - t0)
    T = tr.trel().astype(np.float64)
    o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=3, iters=(10, 10, 10, 0), nselect=8192, mad_bins=256,
                       dtype=np.float64)
    t1 = time.perf_counter()
    o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
    cpu_s = time.perf_counter() - t1
    e = synth.pose_error(T, o.poses()[0])
    et = synth.pose_error(T, Tgt)
    return {"gpu_fps": round(fps, 1), "vs_fp64_oracle_deg": e[0], "vs_fp64_oracle_mm": e[1], "vs_truth_deg": et[0],
            "vs_truth_mm": et[1], "fp64_oracle_cpu_s": round(cpu_s, 3), "note": "host frames in, host pose out (PCIe inclusive)"}


def extra_sequence():
    """300 frames at 640x480 along a smooth trajectory (<= 1.5 deg, <= 20 mm per
    frame), bi-objective lambda 0.5, 61440 photometric points, a keyframe every
    10 frames fused with the next 5: frames/s and drift against the truth."""
    from recon_amd import BIOBJECTIVE, synth, synth_lib
    from recon_amd.hip import DeviceBuffer, hip
    from recon_amd.tracker import PhotoTracker
    W, H, n, seed = 640, 480, 300, 5
    S = synth_lib()
    poses = synth.trajectory(seed * 13 + 5, n, 1.5, 20.0, 0, 1.0)
    fx, fy, cx, cy = synth.pinhole(W, H)
    npx = W * H
    g, d = DeviceBuffer(n * npx), DeviceBuffer(n * npx * 2)
    sc = S.rs_scene_create(ctypes.c_uint64(seed))
    dp = ctypes.POINTER(ctypes.c_double)
    for k in range(n):
        T = np.ascontiguousarray(poses[k], np.float64).ravel()
        S.rs_render_device(sc, T.ctypes.data_as(dp), W, H, fx, fy, cx, cy, ctypes.c_uint64(seed * 1000 + k), 0.01,
                           g.ptr + k * npx, d.ptr + k * npx * 2, None)
    hip().hipDeviceSynchronize()
    S.rs_scene_destroy(sc)
    tr = PhotoTracker(W, H, synth.calib_block_np(W, H), mode=BIOBJECTIVE, n_select=61440, lam=0.5)
    tr.track_sequence(g.ptr, d.ptr, 12, 10, 5)  # warm
    t0 = time.perf_counter()
    P = tr.track_sequence(g.ptr, d.ptr, n, 10, 5)
    dt = time.perf_counter() - t0
    from recon_amd.synth import pose_error
    e = np.array([pose_error(P[k], poses[k]) for k in range(n)])
    return {"frames": n, "fps": round(n / dt, 1), "final_deg": e[-1, 0], "final_mm": e[-1, 1],
            "drift_mm_per_100_frames": e[-1, 1] / (n / 100.0), "max_deg": float(e[:, 0].max()),
            "max_mm": float(e[:, 1].max())}


def pipeline_bench(args):
    """The full pipeline: a seeded 2000-frame sequence at 640x480 with slow (<= 1
    deg, <= 7.5 mm per frame) and fast (<= 8 deg, <= 60 mm per frame) segments of
    100 frames; bi-objective tracking (lambda 0.5, 61440 points) with learned
    warm starts every frame, the plain start as fallback, relocalisation
    against up to 64 keyframes in one batched launch, keyframe fusion (the C
    driver rt_sequence_run). Against the plain tracker (same keyframe policy,
    previous-estimate starts, no relocalisation): frames/s, frames whose motion
    is recovered within 0.5 deg / 5 mm, drift (RMS of the 100-frame relative
    pose error) against ground truth. A step = the whole sequence, frames
    resident in HBM."""
    from recon_amd import synth, synth_lib
    from recon_amd.cnn import PoseNet
    from recon_amd.hip import DeviceBuffer, hip
    from recon_amd.pipeline import SequenceTracker, frame_success, relative_pose_error
    W, H, n, seed = 640, 480, 2000, 61
    S = synth_lib()
    gt = synth.trajectory(seed * 13 + 5, n, 8.0, 60.0, 100, 0.125)
    fx, fy, cx, cy = synth.pinhole(W, H)
    npx = W * H
    g, d = DeviceBuffer(n * npx), DeviceBuffer(n * npx * 2)
    sc = S.rs_scene_create(ctypes.c_uint64(seed))
    dp = ctypes.POINTER(ctypes.c_double)
    for k in range(n):
        T = np.ascontiguousarray(gt[k], np.float64).ravel()
        S.rs_render_device(sc, T.ctypes.data_as(dp), W, H, fx, fy, cx, cy, ctypes.c_uint64(seed * 1000 + k), 0.01,
                           g.ptr + k * npx, d.ptr + k * npx * 2, None)
    hip().hipDeviceSynchronize()
    S.rs_scene_destroy(sc)
    cal = synth.calib_block_np(W, H)
    net = PoseNet.load(args.weights, max_batch=64)
    learned = SequenceTracker(W, H, cal, net=net)
    plain = SequenceTracker(W, H, cal, net=None, relocalise=False)
    learned.run(g, d, 30)  # warm
    plain.run(g, d, 30)
    P, st = learned.run(g, d, n)
    Q, qt = plain.run(g, d, n)
    res = {}
    for name, X, s_ in (("learned", P, st), ("plain", Q, qt)):
        rpe_t, rpe_r = relative_pose_error(X, gt, 100)
        fin = synth.pose_error(X[-1], gt[-1])
        res[name] = {"fps": round(s_["fps"], 1), "frames_recovered": round(frame_success(X, gt), 4),
                     "drift_rpe100_mm": round(rpe_t, 2), "drift_rpe100_deg": round(rpe_r, 3),
                     "final_deg": round(fin[0], 3), "final_mm": round(fin[1], 2), "keyframes": s_["keyframes"],
                     "relocalisations": s_["relocalisations"]}
    vs_plain = synth.pose_error(P[-1], Q[-1])
    log(f"pipeline: learned {res['learned']}; plain {res['plain']}")
    _line("tracked frames/s of the 2000-frame learned pipeline at 640x480", res["learned"]["fps"], "frames/s", args, 1,
          st["seconds"], "f32+f64+bf16",
          {"workload": "2000 frames at 640x480, slow (<= 1 deg, <= 7.5 mm) and fast (<= 8 deg, <= 60 mm per frame) "
                       "segments of 100 frames, bi-objective (lambda 0.5, 61440 points), learned warm starts, plain "
                       "fallback, batched relocalisation against up to 64 keyframes, fusion of 5 frames per keyframe; "
                       "frames resident in HBM", "frames": n, "width": W, "height": H},
          learned=res["learned"], plain=res["plain"],
          final_vs_plain={"deg": vs_plain[0], "mm": vs_plain[1]}, roofline=None, cpu_baseline=None)


def mixed_bench(args):
    """The mixed-size batch: 256 pairs with sizes drawn uniformly from 160x120,
    320x240, 480x360 and 640x480, motions <= 3 deg / 50 mm, 8192 points each,
    4 levels at 640 (else 3), iterations [1, 3, 10, 5], from identity, in one
    persistent launch with the autotuner's variant. A step = keyframe
    preparation + tracking of the batch from frames resident in HBM to poses
    on the host. Roofline of the tracking launch: 60 B per point-iteration
    summed over the pairs (HIP events on the batch's stream). Error per pair
    against the FP64 oracle tracker from the same start, on a sample."""
    import oracle
    from recon_amd import PHOTOMETRIC, synth
    from recon_amd.hip import Event, stream_sync
    from recon_amd.tracker import BatchTracker
    n, N, iters = 256, 8192, (1, 3, 10, 5)
    t0 = time.time()
    pairs, g, d = synth.render_mixed_batch(n, 3, 3.0, 50.0)
    mix = {f"{W}x{H}": sum(1 for p in pairs if (p["width"], p["height"]) == (W, H)) for W, H in synth.MIXED_SIZES}
    log(f"inputs: {n} mixed pairs {mix} rendered in {time.time() - t0:.1f}s")
    bt = BatchTracker(n, 640, 480, mode=PHOTOMETRIC, n_select=N, iters=iters)
    prep = bt.prepare(pairs)
    stream = bt.L.rt_batch_stream(bt.h)
    for _ in range(max(1, args.warmup)):
        Trel, _, rms, st = bt.track(prep)
    steps = max(1, args.steps)
    t1 = time.perf_counter()
    for _ in range(steps):
        Trel, _, rms, st = bt.track(prep)
    step_s = (time.perf_counter() - t1) / steps
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(3):
        bt.L.rt_batch_retrack(bt.h, None, None)
    e1.record(stream)
    e1.synchronize()
    stream_sync(stream)
    k_ms = e0.elapsed_ms(e1) / 3
    alg = sum(sum(iters[:(4 if p["width"] == 640 else 3)]) * N * BYTES_PER_POINT_ITER for p in pairs)
    achieved = alg / (k_ms * 1e-3) / 1e9
    # FP64 oracle from identity on a sample: 8 pairs of each size
    pick = []
    for W, H in synth.MIXED_SIZES:
        pick += [i for i, p in enumerate(pairs) if (p["width"], p["height"]) == (W, H)][:8]
    total = sum(p["width"] * p["height"] for p in pairs)
    gh = g.to_array(np.uint8, 2 * total)
    dh = d.to_array(np.uint16, 2 * total)

    def frames(p):
        W, H = p["width"], p["height"]
        o0, o1 = p["off"]
        return (gh[o0:o0 + W * H].reshape(H, W), dh[o0:o0 + W * H].reshape(H, W),
                gh[o1:o1 + W * H].reshape(H, W), dh[o1:o1 + W * H].reshape(H, W))

    def track_oracle(i, dtype):
        p = pairs[i]
        W, H = p["width"], p["height"]
        o = oracle.Tracker(W, H, p["calib"], mode=oracle.PHOTOMETRIC, levels=4 if W == 640 else 3, iters=iters,
                           nselect=N, dtype=dtype)
        g0, d0, g1, d1 = frames(p)
        o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
        return o.poses()[0]

    from concurrent.futures import ThreadPoolExecutor
    model, nproc = host_cpu()
    with ThreadPoolExecutor(max(1, min(16, nproc))) as ex:
        ref64 = list(ex.map(lambda i: track_oracle(i, np.float64), pick))
    per_pair = []
    for i, R in zip(pick, ref64):
        r, t = synth.pose_error(Trel[i], R)
        rt, tt = synth.pose_error(Trel[i], pairs[i]["gt"])
        per_pair.append({"pair": i, "size": f"{pairs[i]['width']}x{pairs[i]['height']}", "vs_fp64_deg": round(r, 6),
                         "vs_fp64_mm": round(t, 5), "vs_truth_deg": round(rt, 4), "vs_truth_mm": round(tt, 3)})
    ok = [e for e in per_pair if e["vs_truth_deg"] <= 0.1 and e["vs_truth_mm"] <= 2.0]
    cpu = None
    if args.cpu_seconds > 0:
        cycle = [0]

        def work():
            cycle[0] += 1
            track_oracle(pick[cycle[0] % len(pick)], np.float32)
            return 1
        cpu = cpu_sample(work, args.cpu_seconds, args.cpu_threads,
                         f"the {len(pick)} sampled pairs (8 of each size) cycled, FP32 oracle tracker (the reference's "
                         "kernels restated on the CPU), one pair per thread")
    log(f"mixed batch: {n / step_s:.0f} frames/s ({step_s * 1e3:.2f} ms per step); kb_track {k_ms:.3f} ms, "
        f"{achieved:.0f} GB/s algorithmic; status ok {int((st == 0).sum())}/{n}; sample vs FP64 oracle: max "
        f"{max(e['vs_fp64_deg'] for e in ok) if ok else None} deg over {len(ok)} pairs near the truth")
    _line("tracked frames/s of 256 mixed-size pairs (160x120 .. 640x480) in one persistent launch", n / step_s,
          "frames/s", args, steps, step_s, "f32+f64",
          {"workload": "256 pairs, sizes uniform over 160x120, 320x240, 480x360, 640x480 (numpy default_rng(3)), "
                       "rotation <= 3 deg, translation <= 50 mm, 8192 points, iterations [1, 3, 10, 5] (3 levels below "
                       "640 wide), identity start; frames resident in HBM", "pairs": n, "sizes": mix, "points": N},
          status_ok=int((st == 0).sum()),
          pose_error={"sample": len(pick), "near_truth": len(ok), "per_pair": per_pair},
          roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": alg,
                    "kernel": "kb_track (persistent Gauss-Newton), HIP events on the batch's stream",
                    "kernel_ms": round(k_ms, 4)},
          cpu_baseline=cpu)


def fp64_reference(g, d, gt, n, W, H, L, N, iters, cal, threads):
    """The FP64 oracle tracker started at the true motion for the first n
    pairs (the convergence reference): its converged poses, n x 4 x 4."""
    import threading

    import oracle
    npx = W * H
    gh = g.to_array(np.uint8, 2 * n * npx).reshape(2 * n, H, W)
    dh = d.to_array(np.uint16, 2 * n * npx).reshape(2 * n, H, W)
    res = [None] * n
    nt = max(1, threads if threads > 0 else host_cpu()[1])
    nxt = [0]
    lock = threading.Lock()

    def work():
        while True:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= n:
                return
            o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=L, iters=iters, nselect=N, dtype=np.float64)
            o.set_reference(gh[2 * i], dh[2 * i])
            o.set_current(gh[2 * i + 1], dh[2 * i + 1])
            o.prepare()
            o.set_trel(gt[i])
            o.track()
            res[i] = o.poses()[0]

    ts = [threading.Thread(target=work) for _ in range(min(nt, n))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return np.array(res)


if __name__ == "__main__":
    main()
