"""Learned warm start for Richardson-Lucy deconvolution.

A 3-level 3D U-Net (ClarityEx_UNet*: 16/32/64 channels, 3x3x3 convolutions,
ReLU, skip connections, bfloat16 weights on MFMA with float32 accumulation)
maps an observation and its kernel widths to an estimate of a late plain-RL
iterate; a short RL run (ClarityEx_RLPlanReset + ClarityEx_RLPlanIterate)
refines it.  Plain RL from the observation is both the yardstick and the
fallback:

* yardstick -- ``evaluate`` runs plain RL for ``YARDSTICK_ITERATIONS`` on the same
  inputs and measures the refined result against it (relative L2 and
  I-divergence), beside the plain-RL error curve and both wall times;
* fallback -- ``deconvolve`` takes the learned path only when the model
  card stored with the weights says that the yardstick accepted it at the
  requested tolerance and the kernel widths lie in the trained range;
  otherwise it runs plain RL.

Training data (``training_patch``): 128x128x64 patches (z, y, x = 64, 128,
128) with the object statistics of the 256x256x64 benchmark volume scaled
by volume (500 beads, 75 filaments), a Gaussian kernel of the patch's own
size with sigma_xy ~ U[1.2, 2.5] and sigma_z ~ U[2.5, 5.0] (every second
patch also carries the 10 % axial side lobes at z = +-8 of the 512x512x128
benchmark kernel), Poisson noise; the target is the plain-RL estimate after
``TARGET_ITERATIONS`` iterations.  Patch j draws everything from
``numpy.random.default_rng([5, j])``.

The RL semantics are the reference's (src/MaximumLikelihoodDeconvolve.cxx:
44-119, 183-239); the network has no counterpart in the reference.
"""
from __future__ import annotations

import json
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import api, synth

PATCH = (64, 128, 128)            # z, y, x
PATCH_BEADS, PATCH_FILAMENTS = 500, 75
SIGMA_XY = (1.2, 2.5)
SIGMA_Z = (2.5, 5.0)
TARGET_ITERATIONS = 180           # network target
REFINE_ITERATIONS = 20            # RL iterations after the network
YARDSTICK_ITERATIONS = 200        # plain RL that the result is held against
ACCEPT_REL_L2 = 1e-3              # acceptance bar against the yardstick
ACCEPT_IDIV = 1e-3                # relative I-divergence slack

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                       "unet_warmstart.npz")


def log(msg: str) -> None:
    print(f"[learned] {msg}", flush=True)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- patches

@dataclass
class Patch:
    obs: np.ndarray
    psf: np.ndarray
    sigma_xy: float
    sigma_z: float
    side_lobe: float


def patch_object(j: int):
    """Object, kernel and the generator state of training patch j."""
    rng = np.random.default_rng([5, j])
    sxy = float(rng.uniform(*SIGMA_XY))
    sz = float(rng.uniform(*SIGMA_Z))
    lobe = 0.1 if j % 2 else 0.0
    obj = synth.beads_and_filaments(PATCH, PATCH_BEADS, PATCH_FILAMENTS, rng)
    psf = synth.gaussian_kernel(PATCH, sxy, sz, side_lobe=lobe)
    return obj, psf, sxy, sz, lobe, rng


def training_patch(j: int, convolve=None) -> Patch:
    """Training patch j: Poisson(convolve(object, kernel)).  ``convolve``
    defaults to Clarity_Convolve on the GPU."""
    obj, psf, sxy, sz, lobe, rng = patch_object(j)
    conv = convolve or api.Convolve
    obs = synth.poisson(conv(obj, psf), rng)
    return Patch(obs, psf, sxy, sz, lobe)


# ------------------------------------------------------------- model card

def load(path: str = WEIGHTS):
    """(params float32, model card dict) from an .npz written by ``save``
    (no pickles: numpy arrays and a JSON string)."""
    with np.load(path, allow_pickle=False) as z:
        params = np.asarray(z["params"], np.float32)
        card = json.loads(str(z["card"]))
    return params, card


def save(path: str, params, card: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32), card=np.array(json.dumps(card)))


def accepts(card: dict, sigma_xy: float, sigma_z: float,
            tolerance: float = ACCEPT_REL_L2, idiv_tolerance: float = ACCEPT_IDIV) -> bool:
    """Whether the learned path may replace plain RL: the yardstick's measured
    error of the refined result (stored in the card) is within `tolerance`,
    its I-divergence exceeds plain RL's by at most `idiv_tolerance`
    (relative), and the widths are in the trained range."""
    ev = card.get("evaluation") or {}
    err = ev.get("learned_rel_l2")
    idiv_gap = ev.get("learned_idiv_rel_gap")
    if err is None or idiv_gap is None:
        return False
    lo_xy, hi_xy = card["sigma_xy"]
    lo_z, hi_z = card["sigma_z"]
    in_range = lo_xy <= sigma_xy <= hi_xy and lo_z <= sigma_z <= hi_z
    return bool(in_range and err <= tolerance and idiv_gap <= idiv_tolerance)


# --------------------------------------------------------------- training

def make_dataset(first: int, count: int, yardstick: bool = False, workers: int = 8):
    """Observations and plain-RL targets of patches first..first+count-1 on
    the GPU (torch float32 tensors [count, z, y, x]).  With ``yardstick`` the
    plain-RL result after YARDSTICK_ITERATIONS is returned too."""
    import torch
    obs_t = torch.empty((count,) + PATCH, device="cuda")
    tgt_t = torch.empty((count,) + PATCH, device="cuda")
    ref_t = torch.empty((count,) + PATCH, device="cuda") if yardstick else None
    sig = np.zeros((count, 2), np.float32)
    plan = api.RLPlan(PATCH, PATCH)
    t0 = time.time()
    with ThreadPoolExecutor(workers) as pool:
        futs = [pool.submit(training_patch, first + i) for i in range(min(count, 2 * workers))]
        for i in range(count):
            p = futs[i].result()
            if i + 2 * workers < count:
                futs.append(pool.submit(training_patch, first + i + 2 * workers))
            futs[i] = None
            plan.set_inputs(p.obs, p.psf)
            plan.reset()
            plan.iterate(TARGET_ITERATIONS)
            plan.estimate_device(tgt_t[i].data_ptr())
            if yardstick:
                plan.iterate(YARDSTICK_ITERATIONS - TARGET_ITERATIONS)
                plan.estimate_device(ref_t[i].data_ptr())
            obs_t[i].copy_(torch.from_numpy(p.obs))
            sig[i] = (p.sigma_xy, p.sigma_z)
            if (i + 1) % 250 == 0 or i + 1 == count:
                log(f"patches {first}..{first + i}: {time.time() - t0:.1f} s")
    plan.close()
    torch.cuda.synchronize()
    return obs_t, tgt_t, ref_t, sig


def _augment(torch, x, t, code: int):
    """One of the 16 symmetries of the problem: flips of z, y, x and the
    y<->x transpose (Gaussian kernels are symmetric in each axis and equal
    in y and x)."""
    dims = [d for d, bit in ((0, 1), (1, 2), (2, 4)) if code & bit]
    if dims:
        x = torch.flip(x, dims)
        t = torch.flip(t, dims)
    if code & 8:
        x = x.transpose(1, 2)
        t = t.transpose(1, 2)
    return x.contiguous(), t.contiguous()


def predict_rel_l2(net, obs_t, tgt_t, sig) -> float:
    """Mean over patches of the network's relative L2 error to the target."""
    import torch
    out = torch.empty(PATCH, device="cuda")
    errs = []
    for i in range(obs_t.shape[0]):
        net.predict_device(obs_t[i].data_ptr(), PATCH, float(sig[i, 0]), float(sig[i, 1]),
                           out.data_ptr())
        errs.append(float(torch.linalg.vector_norm(out - tgt_t[i]) /
                          torch.linalg.vector_norm(tgt_t[i])))
    return float(np.mean(errs))


def train(n_train: int = 2000, n_val: int = 32, epochs: float = 4.0, lr: float = 1e-3,
          seed: int = 1, init=None, log_every: int = 500):
    """Trains a fresh network (or ``init`` parameters) on the GPU.  Returns
    (UNet, history dict)."""
    import torch
    t0 = time.time()
    obs_t, tgt_t, _, sig = make_dataset(0, n_train)
    vobs, vtgt, _, vsig = make_dataset(n_train, n_val)
    log(f"dataset ready in {time.time() - t0:.1f} s")
    net = api.UNet(seed=seed)
    if init is not None:
        net.set_params(init)
    steps = int(epochs * n_train)
    rng = np.random.default_rng(seed)
    hist = {"steps": steps, "lr": lr, "val_rel_l2": [], "loss": []}
    hist["val_rel_l2"].append((0, predict_rel_l2(net, vobs, vtgt, vsig)))
    log(f"step 0: val rel L2 to RL-{TARGET_ITERATIONS} {hist['val_rel_l2'][-1][1]:.4f}")
    run = 0.0
    t1 = time.time()
    for step in range(1, steps + 1):
        j = int(rng.integers(n_train))
        x, t = _augment(torch, obs_t[j], tgt_t[j], int(rng.integers(16)))
        torch.cuda.synchronize()
        # cosine decay to 1 % of the initial rate
        a = lr * (0.01 + 0.99 * 0.5 * (1 + math.cos(math.pi * step / steps)))
        loss = net.train_step_device(x.data_ptr(), t.data_ptr(), PATCH, float(sig[j, 0]),
                                     float(sig[j, 1]), a)
        run = loss if step == 1 else 0.98 * run + 0.02 * loss
        if step % log_every == 0 or step == steps:
            v = predict_rel_l2(net, vobs, vtgt, vsig)
            hist["val_rel_l2"].append((step, v))
            hist["loss"].append((step, run))
            log(f"step {step}/{steps}: loss {run:.5f} val rel L2 {v:.4f} "
                f"({(time.time() - t1) / step * 1e3:.1f} ms/step)")
    hist["train_seconds"] = time.time() - t1
    del obs_t, tgt_t
    torch.cuda.empty_cache()
    return net, hist


# ------------------------------------------------------------- evaluation

def _timed(fn):
    api.Synchronize()
    t = time.perf_counter()
    fn()
    api.Synchronize()
    return time.perf_counter() - t


def evaluate(net, case=None, sigma_xy: float = 2.0, sigma_z: float = 4.5,
             curve_step: int = 10, repeats: int = 3) -> dict:
    """Measures network + REFINE_ITERATIONS of RL against plain RL after
    YARDSTICK_ITERATIONS on one volume (default: the 512x512x128 benchmark
    volume), with inputs resident on the GPU.  Reports both errors and
    I-divergences, the plain-RL error curve, wall times (median of
    `repeats` after a warm-up) and the time-to-accuracy ratio: plain RL's
    time to reach the learned result's error over the learned path's
    time."""
    import torch
    if case is None:
        case = synth.make_workload("filaments-512x512x128", convolve=api.Convolve)
    shape = case.obs.shape
    obs_d = torch.from_numpy(case.obs).cuda()
    psf_d = torch.from_numpy(np.ascontiguousarray(case.psf, np.float32)).cuda()
    torch.cuda.synchronize()
    plan = api.RLPlan(shape, case.psf.shape)
    plan.set_inputs_device(obs_d.data_ptr(), psf_d.data_ptr())
    ref = torch.empty(shape, device="cuda")
    cur = torch.empty(shape, device="cuda")
    warm = torch.empty(shape, device="cuda")

    def plain(k):
        plan.reset()
        plan.iterate(k)

    def learned():
        net.predict_device(obs_d.data_ptr(), shape, sigma_xy, sigma_z, warm.data_ptr())
        plan.reset_device(warm.data_ptr())
        plan.iterate(REFINE_ITERATIONS)

    def med(fn):
        fn()
        return float(np.median([_timed(fn) for _ in range(repeats)]))

    t_plain = med(lambda: plain(YARDSTICK_ITERATIONS))
    t_learned = med(learned)
    t_net = med(lambda: net.predict_device(obs_d.data_ptr(), shape, sigma_xy, sigma_z,
                                           warm.data_ptr()))
    plain(YARDSTICK_ITERATIONS)
    plan.estimate_device(ref.data_ptr())
    nref = float(torch.linalg.vector_norm(ref))

    def err(x):
        return float(torch.linalg.vector_norm(x - ref)) / nref

    curve = []
    plan.reset()
    done = 0
    for k in range(curve_step, YARDSTICK_ITERATIONS + 1, curve_step):
        plan.iterate(k - done)
        done = k
        plan.estimate_device(cur.data_ptr())
        curve.append((k, err(cur)))
    # at the stated tolerance (ACCEPT_REL_L2 of the yardstick): the first
    # plain-RL iteration count that reaches it (one by one near the end, where
    # the curve crosses it), and the first refinement count after the
    # network that reaches it (every 10 iterations, up to the yardstick's
    # count)
    tol = ACCEPT_REL_L2
    k_plain_tol = None
    start = max(0, YARDSTICK_ITERATIONS - 60)
    for k, e in curve:
        if e <= tol:
            start = max(0, k - curve_step)
            break
    plan.reset()
    plan.iterate(start)
    for k in range(start + 1, YARDSTICK_ITERATIONS + 1):
        plan.iterate(1)
        plan.estimate_device(cur.data_ptr())
        if err(cur) <= tol:
            k_plain_tol = k
            break
    net.predict_device(obs_d.data_ptr(), shape, sigma_xy, sigma_z, warm.data_ptr())
    plan.reset_device(warm.data_ptr())
    k_learned_tol, best_learned, learned_curve = None, float("inf"), []
    done = 0
    for k in range(curve_step, YARDSTICK_ITERATIONS + 1, curve_step):
        plan.iterate(k - done)
        done = k
        plan.estimate_device(cur.data_ptr())
        e = err(cur)
        learned_curve.append((k, e))
        best_learned = min(best_learned, e)
        if e <= tol:
            k_learned_tol = k
            break
    learned()
    plan.estimate_device(cur.data_ptr())
    e_learned = err(cur)
    obs_np = case.obs
    est_l = cur.cpu().numpy()
    est_p = ref.cpu().numpy()
    idiv_l = api.IDivergence(obs_np, api.Convolve(est_l, case.psf))
    idiv_p = api.IDivergence(obs_np, api.Convolve(est_p, case.psf))
    e_warm = err(warm)
    # plain RL iterations (linear interpolation on the curve) and time to
    # reach the learned path's error
    k_eq = None
    prev = (0, 1.0)
    for k, e in curve:
        if e <= e_learned:
            k0, e0 = prev
            k_eq = k0 + (k - k0) * (e0 - e_learned) / max(e0 - e, 1e-30)
            break
        prev = (k, e)
    per_it = t_plain / YARDSTICK_ITERATIONS
    t_eq = None if k_eq is None else k_eq * per_it
    t_plain_tol = None if k_plain_tol is None else k_plain_tol * per_it
    t_learned_tol = None if k_learned_tol is None else t_net + k_learned_tol * per_it
    ev = {
        "volume": list(shape),
        "sigma_xy": sigma_xy, "sigma_z": sigma_z,
        "refine_iterations": REFINE_ITERATIONS,
        "yardstick_iterations": YARDSTICK_ITERATIONS,
        "learned_rel_l2": e_learned,
        "network_rel_l2": e_warm,
        "plain_curve": curve,
        "idiv_plain": idiv_p, "idiv_learned": idiv_l,
        "learned_idiv_rel_gap": (idiv_l - idiv_p) / idiv_p,
        "seconds_plain": t_plain, "seconds_learned": t_learned, "seconds_network": t_net,
        "wall_ratio_vs_yardstick": t_plain / t_learned,
        "plain_iterations_to_same_error": k_eq,
        "time_to_accuracy_speedup": None if t_eq is None else t_eq / t_learned,
        "tolerance": tol,
        "plain_iterations_to_tolerance": k_plain_tol,
        "plain_seconds_to_tolerance": t_plain_tol,
        "learned_refine_iterations_to_tolerance": k_learned_tol,
        "learned_seconds_to_tolerance": t_learned_tol,
        "learned_best_rel_l2_within_yardstick_iterations": best_learned,
        "learned_curve": learned_curve,
        "speedup_at_tolerance": (None if t_plain_tol is None or t_learned_tol is None
                                 else t_plain_tol / t_learned_tol),
        "accepted": bool(e_learned <= ACCEPT_REL_L2 and
                         (idiv_l - idiv_p) / idiv_p <= ACCEPT_IDIV),
    }
    plan.close()
    return ev


# ------------------------------------------------------------ deconvolve

def deconvolve(obs, psf, sigma_xy: float, sigma_z: float, iterations: int = YARDSTICK_ITERATIONS,
               tolerance: float = ACCEPT_REL_L2, idiv_tolerance: float = ACCEPT_IDIV,
               weights: str = WEIGHTS):
    """Deconvolves `obs` (z, y, x float32, host) with `psf`.  Takes the
    learned path (network + REFINE_ITERATIONS of RL) when the model card
    accepts it at `tolerance` / `idiv_tolerance` for these widths and `iterations` is the
    yardstick's count; otherwise plain RL for `iterations` (the fallback).
    Returns (estimate, path) with path "learned" or "plain"."""
    obs = np.ascontiguousarray(obs, np.float32)
    card = None
    if os.path.exists(weights):
        params, card = load(weights)
    use = (card is not None and iterations == card.get("yardstick_iterations") and
           accepts(card, sigma_xy, sigma_z, tolerance, idiv_tolerance) and
           all(n % 4 == 0 for n in obs.shape))
    plan = api.RLPlan(obs.shape, psf.shape)
    try:
        plan.set_inputs(obs, psf)
        if use:
            net = api.UNet()
            try:
                net.set_params(params)
                plan.reset(net.predict(obs, sigma_xy, sigma_z))
            finally:
                net.close()
            plan.iterate(card["refine_iterations"])
        else:
            plan.reset()
            plan.iterate(iterations)
        return plan.estimate(), ("learned" if use else "plain")
    finally:
        plan.close()


# ------------------------------------------------------------------- main

def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train / evaluate the learned warm start")
    ap.add_argument("--train", type=int, default=2000, help="training patches")
    ap.add_argument("--epochs", type=float, default=4.0)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--eval-only", default=None, help="evaluate these weights (.npz)")
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args(argv)
    api.Register()
    log(f"device {api.DeviceName()}")
    if a.eval_only:
        params, card = load(a.eval_only)
        net = api.UNet()
        net.set_params(params)
        hist = card.get("training", {})
    else:
        init = load(a.init)[0] if a.init else None
        net, hist = train(a.train, epochs=a.epochs, lr=a.lr, init=init)
        card = {}
    ev = evaluate(net)
    log("evaluation " + json.dumps({k: v for k, v in ev.items() if k != "plain_curve"}))
    card.update({
        "patch": list(PATCH), "n_train": a.train, "target_iterations": TARGET_ITERATIONS,
        "refine_iterations": REFINE_ITERATIONS, "yardstick_iterations": YARDSTICK_ITERATIONS,
        "sigma_xy": list(SIGMA_XY), "sigma_z": list(SIGMA_Z), "training": hist,
        "evaluation": ev,
    })
    save(a.out, net.get_params(), card)
    log(f"wrote {a.out}")
    net.close()
    api.UnRegister()


if __name__ == "__main__":
    main()
