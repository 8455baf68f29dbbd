"""Richardson-Lucy throughput on one MI355X.

Workload: plain Richardson-Lucy (Clarity_MaximumLikelihoodDeconvolve
semantics), 100 iterations on a float32 512x512x128 volume with a same-size
kernel (FFT work domain 1024x1024x256): 16000 beads + 2400 random-walk
filaments on background 10, anisotropic Gaussian kernel (sigma_xy 2.0,
sigma_z 4.5, 10 % axial side lobes), Poisson noise -- synthetic, seeded.

One step = one full deconvolution with the observation and kernel already
resident in HBM: pad + kernel transform + energy, 100 RL iterations, clip
into a device output buffer.  Metric: image voxels x iterations / s / 1e6.

Also reported:
  roofline      for the dominant kernel, from HIP events recorded on the
                library stream around each launch, over a second pass of
                the same timed steps (the first pass, which gives `value`,
                replays the plan's HIP graphs without per-launch events:
                recording them costs ~4 % of the iteration time);
                algorithmic bytes per launch from the plan's own model.
  cpu_baseline  the CPU oracle (the reference's CPU algorithm restated in C,
                FFTW3-API FFTs via MKL when present, plan per FFT), 10
                iterations of the same volume on this host's OpenMP threads,
                both reduction modes, deconvolution and whole-call scopes,
                median of 3 after a warm-up; value = serial float32 sum
                (the reference's), deconvolution scope.
  learned_warm_start
                the metric's second half: the committed network + 20 RL
                iterations against plain RL on the same volume, measured against
                plain RL after 200 iterations (relative L2, I-divergence),
                with plain RL's time to reach the same error.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
    python bench.py --suite      # every other driver on the path, one line each
    python bench.py --network    # the learned warm start's network alone
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

try:
    # PyTorch-ROCm ships its own HIP runtime under /opt/rocm's soname; loading
    # it first keeps one runtime in the process (libclarity binds to it)
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime calls for device buffers (plumbing only)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]

    def malloc(self, n):
        p = ctypes.c_void_p()
        rc = self.L.hipMalloc(ctypes.byref(p), n)
        if rc != 0:
            raise RuntimeError(f"hipMalloc({n}) failed: {rc}")
        return p.value

    def h2d(self, dst, arr):
        rc = self.L.hipMemcpy(ctypes.c_void_p(dst), arr.ctypes.data_as(ctypes.c_void_p),
                              arr.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def d2h(self, arr, src):
        rc = self.L.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src),
                              arr.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")

    def d2d(self, dst, src, n):
        rc = self.L.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), n, 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2D failed: {rc}")

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))


def make_inputs(api, shape, seed=3):
    """The filament workload's statistics (2000 beads and 300 filaments per
    256x256x64, scaled by volume) on `shape`; observation convolved on the
    GPU."""
    from clarity_amd import synth
    rng = np.random.default_rng(seed)
    base = (64, 256, 256)
    f = np.prod(shape) / np.prod(base)
    obj = synth.beads_and_filaments(shape, int(round(2000 * f)), int(round(300 * f)), rng)
    psf = synth.gaussian_kernel(shape, 2.0, 4.5, side_lobe=0.1)
    blurred = api.Convolve(obj, psf)
    obs = synth.poisson(blurred, rng)
    return obs, psf


def host_cpu():
    """lscpu's model, sockets, cores per socket, threads per core and the
    logical CPU count of this host."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets",
            "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core",
            "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys and keys[k.strip()] not in info:
            info[keys[k.strip()]] = v.strip()
    return info


def cpu_baseline(obs, psf, iterations, threads):
    """The oracle's RL on this host (the reference's CPU algorithm: FFTW3 API
    with a plan per call, a buffer per convolution), `iterations`
    iterations of the full volume, in both reduction modes (the reference's
    sequential float32 energy sum, and float64).  Per mode one warm-up call
    of one iteration, then the median of 3 calls, in the reference's two
    timing scopes: deconvolution only (TIME_DECONVOLVE, after padding and
    before clipping) and the whole call (TIME_TOTAL)."""
    import oracle
    backend = "own-fp64"
    if oracle.set_fft_backend(oracle.FFT_FFTW_API):
        backend = oracle.fft_backend_name()
    oracle.set_threads(threads)
    res = {}
    for mode, name in ((oracle.REDUCE_SERIAL_F32, "serial_f32"), (oracle.REDUCE_FP64, "fp64")):
        oracle.set_reduce_mode(mode)
        oracle.ml(obs, psf, 1)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.ml(obs, psf, iterations)
            total = time.perf_counter() - t0
            runs.append((oracle.last_deconvolve_seconds(), total, oracle.last_loop_seconds()))
            log(f"[bench] cpu {name}: deconvolve {runs[-1][0]:.2f} s, total {total:.2f} s")
        med = [float(np.median([r[i] for r in runs])) for i in range(3)]
        res[name] = {"deconvolve_s": round(med[0], 3), "total_s": round(med[1], 3),
                     "iterations_s": round(med[2], 3)}
    oracle.set_reduce_mode(oracle.REDUCE_FP64)
    oracle.set_fft_backend(oracle.FFT_OWN)
    return res, backend


def dense_bytes(shape, passes_real, passes_complex):
    """Bytes of `passes_real` real and `passes_complex` half-spectrum sweeps of
    the work domain of an image `shape` with a same-size kernel."""
    nz, ny, nx = (2 * s for s in shape)
    return 4.0 * nx * ny * nz * passes_real + 8.0 * (nx // 2 + 1) * ny * nz * passes_complex


def network_only(api, reps=7):
    """The learned warm start's network alone: U-Net inference on the
    512x512x128 filament volume (sigma_xy 2.0, sigma_z 4.5), observation
    and estimate in device memory; median of `reps` after two warm-ups.
    One JSON line (the per-kernel profile of the network runs this)."""
    from clarity_amd import learned as L
    shape = (128, 512, 512)
    obs, _ = make_inputs(api, shape)
    hip = Hip()
    d_obs, d_out = hip.malloc(obs.nbytes), hip.malloc(obs.nbytes)
    hip.h2d(d_obs, obs)
    params, _ = L.load()
    net = api.UNet()
    net.set_params(params)
    ts = []
    for i in range(reps + 2):
        api.Synchronize()
        t = time.perf_counter()
        net.predict_device(d_obs, shape, 2.0, 4.5, d_out)
        api.Synchronize()
        if i >= 2:
            ts.append(time.perf_counter() - t)
    net.close()
    hip.free(d_obs)
    hip.free(d_out)
    print(json.dumps({"network": "learned warm start U-Net inference",
                      "workload": "filaments 512x512x128", "reps": reps,
                      "ms_median": round(1e3 * float(np.median(ts)), 3),
                      "ms_min": round(1e3 * float(np.min(ts)), 3)}), flush=True)


def suite(api, threads):
    """Every other driver on the path, GPU vs the CPU oracle on the same
    inputs (a 256x256x64 volume with a same-size kernel; batched RL on 48
    uneven volumes of 64..256 x 64..256 x 32..96).  GPU times are whole C-API calls with
    host buffers (upload, pad, compute, clip, download); the dense-model
    bandwidth counts only the work-domain sweeps of the algorithm."""
    import oracle
    from clarity_amd import synth
    oracle.set_fft_backend(oracle.FFT_FFTW_API)
    oracle.set_reduce_mode(oracle.REDUCE_SERIAL_F32)
    oracle.set_threads(threads)
    backend = oracle.fft_backend_name()
    shape = (64, 256, 256)
    rng = np.random.default_rng(2)
    obj = synth.beads_and_filaments(shape, 2000, 300, rng)
    psf = synth.gaussian_kernel(shape, 2.0, 4.5, side_lobe=0.1)
    obs = synth.poisson(api.Convolve(obj, psf), rng)

    def gpu_time(fn, reps=3):
        # the CPU oracle's OpenMP/MKL threads keep spinning for a while after
        # their last parallel region and would compete with the library's
        # host-side copies: let them go idle first
        time.sleep(0.5)
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]  # median

    def cpu_time(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    def emit(name, g, c, byts, note=""):
        line = {"suite": name, "gpu_ms": round(g * 1e3, 3), "cpu_ms": round(c * 1e3, 1),
                "speedup": round(c / g, 1), "cpu_kind": "port", "cpu_fft": backend,
                "cores": threads}
        if byts:
            line["dense_GBps"] = round(byts / g / 1e9, 1)
        if note:
            line["note"] = note
        print(json.dumps(line), flush=True)

    emit("convolve 256x256x64", gpu_time(lambda: api.Convolve(obj, psf)),
         cpu_time(lambda: oracle.convolve(obj, psf)), dense_bytes(shape, 3, 14))
    emit("wiener eps=1e-3 256x256x64", gpu_time(lambda: api.WienerDeconvolve(obs, psf, 1e-3)),
         cpu_time(lambda: oracle.wiener(obs, psf, 1e-3)), dense_bytes(shape, 3, 14))
    g = gpu_time(lambda: api.JansenVanCittertDeconvolve(obs, psf, 50), reps=1)
    c1 = cpu_time(lambda: oracle.jvc(obs, psf, 1))
    c3 = cpu_time(lambda: oracle.jvc(obs, psf, 3))
    emit("jvc 50 iterations 256x256x64", g, c1 + 49 * max(c3 - c1, 0.0) / 2,
         dense_bytes(shape, 4 * 50, 9 * 50),
         "cpu: 1-iteration call + 49 x per-iteration cost from a 3-iteration call")
    g = gpu_time(lambda: api.MaximumLikelihoodDeconvolve(obs, psf, 100), reps=1)
    oracle.ml(obs, psf, 2)
    loop = oracle.last_loop_seconds() / 2
    emit("rl 100 iterations 256x256x64 (C API, host buffers)", g, 100 * loop,
         dense_bytes(shape, 4 * 100, 18 * 100), "cpu: 100 x per-iteration loop time of 2")
    k0 = synth.gaussian_kernel(shape, 2.4, 5.4, side_lobe=0.1)
    g = gpu_time(lambda: api.BlindMaximumLikelihoodDeconvolve(obs, k0, 100), reps=1)
    cb = cpu_time(lambda: oracle.blind_ml(obs, k0, 1))
    emit("blind ml as written 100 iterations 256x256x64", g, 100 * cb, 0,
         "cpu: 100 x the cost of a 1-iteration call")
    g = gpu_time(lambda: api.BlindRL(obs, k0, 4, 5), reps=1)
    cb = cpu_time(lambda: oracle.blind_rl(obs, k0, 1, 1))
    emit("blind rl 4 outer x 5 inner 256x256x64", g, 20 * cb, 0,
         "cpu: 20 x the cost of a 1 x 1 call")
    # the 512x512x128 volume's blind run: 20 outer x 5 inner, starting from
    # the kernel's Gaussian 20 % too wide
    big = (128, 512, 512)
    rng = np.random.default_rng(3)
    obj_b = synth.beads_and_filaments(big, 16000, 2400, rng)
    psf_b = synth.gaussian_kernel(big, 2.0, 4.5, side_lobe=0.1)
    obs_b = synth.poisson(api.Convolve(obj_b, psf_b), rng)
    k0_b = synth.gaussian_kernel(big, 2.4, 5.4, side_lobe=0.1)
    g = gpu_time(lambda: api.BlindRL(obs_b, k0_b, 20, 5), reps=1)
    cb = cpu_time(lambda: oracle.blind_rl(obs_b, k0_b, 1, 1))
    emit("blind rl 20 outer x 5 inner 512x512x128", g, 100 * cb, 0,
         "cpu: 100 x the cost of a 1 x 1 call")
    del obj_b, psf_b, obs_b, k0_b
    # 48 volumes, x,y in {64..256}, z in {32..96}, own Gaussian kernels
    imgs, ks = synth.make_batch(convolve=api.Convolve)
    vox = sum(int(np.prod(a.shape)) for a in imgs)
    g = gpu_time(lambda: api.MaximumLikelihoodDeconvolveBatch(imgs, ks, 50))
    # the same call with no iterations: plans, uploads, kernel transforms,
    # downloads; the difference is the 50 grouped iterations alone
    g0 = gpu_time(lambda: api.MaximumLikelihoodDeconvolveBatch(imgs, ks, 0))
    t_cpu = 0.0
    for a, k in zip(imgs[:3], ks[:3]):
        oracle.ml(a, k, 1)
        t_cpu += oracle.last_loop_seconds() * vox / sum(int(np.prod(x.shape)) for x in imgs[:3])
    # bytes the volumes' kernels move per iteration (each plan's per-launch
    # model x launches per iteration)
    moved = 0.0
    for a, k in zip(imgs, ks):
        p = api.RLPlan(a.shape, k.shape)
        moved += sum(b * n for _, b, n in p.describe())
        p.close()
    line = {"suite": "batched rl 48 uneven volumes x 50 iterations",
            "gpu_ms": round(g * 1e3, 2), "Mvoxel_iters_per_s": round(vox * 50 / g / 1e6, 1),
            "setup_transfer_ms": round(g0 * 1e3, 2),
            "iterations_Mvoxel_iters_per_s": round(vox * 50 / max(g - g0, 1e-9) / 1e6, 1),
            "moved_bytes_per_iteration": moved,
            "iterations_moved_GBps": round(moved * 50 / max(g - g0, 1e-9) / 1e9, 1),
            "cpu_Mvoxel_iters_per_s": round(vox / t_cpu / 1e6, 2), "voxels": vox,
            "cpu_note": "1 iteration on 3 of the volumes, scaled by voxels",
            "cpu_fft": backend, "cores": threads}
    print(json.dumps(line), flush=True)
    oracle.set_reduce_mode(oracle.REDUCE_FP64)
    oracle.set_fft_backend(oracle.FFT_OWN)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iterations", type=int, default=100)
    ap.add_argument("--shape", type=str, default="128,512,512", help="image z,y,x")
    ap.add_argument("--cpu-iterations", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm start's time-to-accuracy measurement")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device 3 times "
                         "(a plain copy kernel of known size, to calibrate rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--suite", action="store_true",
                    help="instead of the headline line, time every other driver on "
                         "the path (GPU vs CPU oracle), one JSON line each")
    ap.add_argument("--network", action="store_true",
                    help="instead of the headline line, time the learned warm start's "
                         "network alone (U-Net inference on the 512x512x128 volume)")
    args = ap.parse_args()

    if args.network:
        from clarity_amd import api
        api.Register()
        network_only(api)
        api.UnRegister()
        return

    if args.suite:
        from clarity_amd import api
        api.Register()
        suite(api, min(16, os.cpu_count() or 1))
        api.UnRegister()
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this "
              f"build (each box has one MI355X; the multi-GPU sharding is design-only).",
              file=sys.stderr)
        sys.exit(2)

    from clarity_amd import api
    shape = tuple(int(v) for v in args.shape.split(","))
    api.Register()
    log(f"[bench] device {api.DeviceName()}; image {shape}; building inputs")
    obs, psf = make_inputs(api, shape)
    hip = Hip()
    d_obs = hip.malloc(obs.nbytes)
    d_psf = hip.malloc(psf.nbytes)
    d_out = hip.malloc(obs.nbytes)
    hip.h2d(d_obs, obs)
    hip.h2d(d_psf, psf)

    plan = api.RLPlan(shape, psf.shape)
    log(f"[bench] engine {plan.engine}")

    def step():
        plan.set_inputs_device(d_obs, d_psf)
        plan.iterate(args.iterations)
        plan.estimate_device(d_out)

    for i in range(args.warmup):
        step()
        api.Synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")

    api.Synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    api.Synchronize()
    elapsed = time.perf_counter() - t0
    log(f"[bench] {args.steps} steps in {elapsed:.3f} s")

    vox = int(np.prod(shape))
    value = vox * args.iterations * args.steps / elapsed / 1e6
    ms_per_step = 1e3 * elapsed / args.steps

    # roofline of the dominant kernel: the same steps again, with HIP events
    # around every launch on the library stream
    roof = None
    kernels = plan.describe()
    if kernels:
        plan.set_timing(True)
        plan.timings(reset=True)
        t1 = time.perf_counter()
        for i in range(args.steps):
            step()
        api.Synchronize()
        log(f"[bench] {args.steps} steps with per-launch events in "
            f"{time.perf_counter() - t1:.3f} s")
        tm = plan.timings()
        rows = []
        for (name, byts, per_it), (cnt, ms) in zip(kernels, tm):
            if cnt:
                rows.append((ms, name, byts, cnt))
        rows.sort(reverse=True)
        ms, name, byts, cnt = rows[0]
        avg_s = ms / cnt / 1e3
        achieved = byts / avg_s / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tf):
            traffic = json.load(open(tf)).get("bytes_per_launch", {}).get(name)
        roof = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": byts,
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": cnt,
                "timing": "HIP events on the library stream around each launch, "
                          "over a second pass of the timed steps"}
        for ms_, name_, byts_, cnt_ in rows:
            log(f"[bench] {name_:24s} launches {cnt_:6d} avg {ms_ / cnt_ * 1e3:9.2f} us "
                f"{byts_ / (ms_ / cnt_ / 1e3) / 1e9:8.1f} GB/s")
        plan.set_timing(False)

    # whole-iteration rates against three byte counts per iteration: the
    # dense model 4R + 18C (nine passes over the whole reference work
    # domain, image + kernel), the pruned model 0.5R + 8C (the same passes
    # over the nonzero corner only, with a stored OTF), both on the
    # reference domain, and the bytes this plan's kernels actually move
    # (its per-launch model x launches per iteration, on its own, smaller
    # alias-free domain: the pruned fraction can exceed 1, the moved one
    # cannot)
    nz, ny, nx = (2 * s for s in shape)
    R = 4.0 * nx * ny * nz
    C = 8.0 * (nx // 2 + 1) * ny * nz
    it_s = elapsed / (args.steps * args.iterations)
    moved = sum(b * n for _, b, n in kernels) if kernels else None

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        hostinfo = host_cpu()
        log(f"[bench] cpu baseline: {args.cpu_iterations} iterations on {threads} threads "
            f"({hostinfo})")
        res, backend = cpu_baseline(obs, psf, args.cpu_iterations, threads)
        rate = {k: {"deconvolve": round(vox * args.cpu_iterations / v["deconvolve_s"] / 1e6, 3),
                    "total": round(vox * args.cpu_iterations / v["total_s"] / 1e6, 3),
                    "iterations_only": round(vox * args.cpu_iterations / v["iterations_s"] / 1e6,
                                             3)}
                for k, v in res.items()}
        cpu = {"value": rate["serial_f32"]["deconvolve"],
               "unit": "Mvoxel-iters/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_iterations} RL iterations of the same "
                         f"{shape[2]}x{shape[1]}x{shape[0]} volume (work "
                         f"{nx}x{ny}x{nz}); value = the reference's arithmetic (serial "
                         f"float32 energy sum), deconvolution scope (TIME_DECONVOLVE: "
                         f"kernel transform, energy, iterations), median of 3 after one "
                         f"warm-up call; FFT {backend} on {threads} threads",
               "Mvoxel_iters_per_s": rate, "seconds": res, "host": hostinfo,
               "fft": backend}
        log(f"[bench] cpu baseline {cpu['value']} Mvoxel-iters/s; {rate}")

    # the whole C-API call with host buffers (upload, pad, kernel transform,
    # iterations, clip, download: the reference's TIME_TOTAL scope), beside
    # `value`'s resident deconvolution
    api.MaximumLikelihoodDeconvolve(obs, psf, args.iterations)
    tc = []
    for _ in range(3):
        t1 = time.perf_counter()
        api.MaximumLikelihoodDeconvolve(obs, psf, args.iterations)
        tc.append(time.perf_counter() - t1)
    c_api_ms = 1e3 * float(np.median(tc))
    log(f"[bench] C-API call with host buffers: {c_api_ms:.1f} ms")

    # time-to-accuracy of the learned warm start against plain RL, on the
    # same volume (network + 20 RL iterations vs plain RL, measured against
    # plain RL after 200 iterations)
    learned = None
    if not args.no_learned and shape == (128, 512, 512):
        from clarity_amd import learned as L
        if os.path.exists(L.WEIGHTS):
            log("[bench] learned warm start: network + 20 RL vs plain RL-200")
            params, card = L.load()
            net = api.UNet()
            net.set_params(params)
            ev = L.evaluate(net, case=type("Case", (), {"obs": obs, "psf": psf})())
            net.close()
            learned = {
                "learned_ms": round(ev["seconds_learned"] * 1e3, 2),
                "network_ms": round(ev["seconds_network"] * 1e3, 2),
                "plain_rl200_ms": round(ev["seconds_plain"] * 1e3, 2),
                "learned_rel_l2_vs_rl200": round(ev["learned_rel_l2"], 4),
                "learned_idiv_rel_gap": round(ev["learned_idiv_rel_gap"], 5),
                "plain_iterations_to_same_error": None if ev["plain_iterations_to_same_error"]
                is None else round(ev["plain_iterations_to_same_error"], 1),
                "time_to_accuracy_speedup": None if ev["time_to_accuracy_speedup"] is None
                else round(ev["time_to_accuracy_speedup"], 2),
                "wall_ratio_vs_rl200": round(ev["wall_ratio_vs_yardstick"], 2),
                "accepted_at_1e-3": ev["accepted"],
                "at_1e-3": {
                    "plain_iterations": ev["plain_iterations_to_tolerance"],
                    "plain_ms": None if ev["plain_seconds_to_tolerance"] is None
                    else round(ev["plain_seconds_to_tolerance"] * 1e3, 2),
                    "learned_refine_iterations": ev["learned_refine_iterations_to_tolerance"],
                    "learned_ms": None if ev["learned_seconds_to_tolerance"] is None
                    else round(ev["learned_seconds_to_tolerance"] * 1e3, 2),
                    "learned_best_rel_l2": round(
                        ev["learned_best_rel_l2_within_yardstick_iterations"], 5),
                    "speedup": None if ev["speedup_at_tolerance"] is None
                    else round(ev["speedup_at_tolerance"], 2),
                },
            }
            log(f"[bench] learned {learned}")

    line = {
        "metric": "Mvoxel-iters/s RL 512x512x128 fp32; time-to-accuracy learned vs plain RL",
        "value": round(value, 2),
        "unit": "Mvoxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"richardson-lucy {args.iterations} iterations, image "
                               f"{shape[2]}x{shape[1]}x{shape[0]}, same-size kernel, "
                               f"reference work domain {nx}x{ny}x{nz}",
                   "engine": plan.engine,
                   "engine_work_domain": "x".join(str(v) for v in plan.work_shape[::-1]),
                   "iterations": args.iterations,
                   "ms_per_iteration": round(1e3 * it_s, 4),
                   "dense_model_GBps": round((4 * R + 18 * C) / it_s / 1e9, 1),
                   "pruned_model_GBps": round((0.5 * R + 8 * C) / it_s / 1e9, 1),
                   "pruned_model_frac": round((0.5 * R + 8 * C) / it_s / 1e9 / HBM_PEAK_GBS, 4),
                   "moved_bytes_per_iteration": moved,
                   "moved_GBps": round(moved / it_s / 1e9, 1) if moved else None,
                   "moved_frac": round(moved / it_s / 1e9 / HBM_PEAK_GBS, 4) if moved else None,
                   "c_api_call_ms": round(c_api_ms, 2),
                   "c_api_Mvoxel_iters_per_s": round(vox * args.iterations / c_api_ms * 1e3 / 1e6,
                                                     1)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "learned_warm_start": learned,
    }
    print(json.dumps(line), flush=True)
    if args.copy_mib:
        # a plain copy kernel of known size (torch's device copy, which runs
        # the HIP runtime's copyBuffer kernel): reads and writes exactly n
        # bytes, so counter readings scale against it
        import torch
        n = args.copy_mib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"[bench] copied {args.copy_mib} MiB device-to-device 3 times (torch copy kernel)")
    plan.close()
    for p in (d_obs, d_psf, d_out):
        hip.free(p)
    api.UnRegister()


if __name__ == "__main__":
    main()
