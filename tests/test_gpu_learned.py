"""Learned warm start on the GPU, measured against plain RL: the committed network
(clarity_amd/data/unet_warmstart.npz, trained by clarity_amd.learned on
seeded patches 0..1999, validated on 2000..2031) on held-out patches, and
the gate of clarity_amd.learned.deconvolve between the learned path and
plain RL."""
import os

import numpy as np
import pytest

from clarity_amd import learned as L

pytestmark = pytest.mark.gpu

HELD_OUT = [100000, 100001, 100002, 100003]   # never drawn for training


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def trained(clarity_gpu):
    if not os.path.exists(L.WEIGHTS):
        pytest.skip("no trained weights committed")
    params, card = L.load()
    net = clarity_gpu.UNet()
    net.set_params(params)
    yield net, card
    net.close()


def test_gain_over_plain_rl(clarity_gpu, trained):
    """Network + 20 RL iterations ends closer to RL-200 than plain RL after
    20 iterations (the same refinement without the network) and after 40
    (plain RL given more than the network's cost: at this size the network
    costs about as much as 12 RL iterations)."""
    api = clarity_gpu
    net, _ = trained
    plan = api.RLPlan(L.PATCH, L.PATCH)
    rows = []
    for j in HELD_OUT:
        p = L.training_patch(j)
        print(f"[learned] patch {j} sigma {p.sigma_xy:.2f}/{p.sigma_z:.2f}", flush=True)
        plan.set_inputs(p.obs, p.psf)
        plan.reset()
        plan.iterate(20)
        plain20 = plan.estimate()
        plan.iterate(20)
        plain40 = plan.estimate()
        plan.iterate(L.YARDSTICK_ITERATIONS - 40)
        yard = plan.estimate()
        plan.reset(net.predict(p.obs, p.sigma_xy, p.sigma_z))
        plan.iterate(L.REFINE_ITERATIONS)
        learned = plan.estimate()
        rows.append((rel_l2(learned, yard), rel_l2(plain20, yard), rel_l2(plain40, yard)))
        print(f"[learned]   rel L2 to RL-200: learned {rows[-1][0]:.4f} "
              f"plain-20 {rows[-1][1]:.4f} plain-40 {rows[-1][2]:.4f}", flush=True)
    plan.close()
    for e_l, e20, e40 in rows:
        assert e_l < 0.5 * e20
        assert e_l < e40


def test_gate_falls_back_to_plain_rl(clarity_gpu, trained):
    api = clarity_gpu
    net, card = trained
    p = L.training_patch(HELD_OUT[0])
    plan = api.RLPlan(L.PATCH, L.PATCH)
    plan.set_inputs(p.obs, p.psf)
    plan.reset()
    plan.iterate(L.YARDSTICK_ITERATIONS)
    plain = plan.estimate()
    plan.reset(net.predict(p.obs, p.sigma_xy, p.sigma_z))
    plan.iterate(L.REFINE_ITERATIONS)
    learned = plan.estimate()
    plan.close()
    ev = card["evaluation"]
    # the card's own decision at the acceptance bar
    est, path = L.deconvolve(p.obs, p.psf, p.sigma_xy, p.sigma_z)
    assert path == ("learned" if ev["accepted"] else "plain")
    if path == "plain":
        assert np.array_equal(est, plain)
    # widths outside the trained range always take plain RL
    est, path = L.deconvolve(p.obs, p.psf, 3.5, p.sigma_z, tolerance=1.0, idiv_tolerance=1.0)
    assert path == "plain" and np.array_equal(est, plain)
    # a caller who accepts the measured error gets the learned path
    est, path = L.deconvolve(p.obs, p.psf, p.sigma_xy, p.sigma_z,
                             tolerance=1.5 * ev["learned_rel_l2"],
                             idiv_tolerance=max(1.5 * ev["learned_idiv_rel_gap"], 1e-3))
    assert path == "learned" and np.array_equal(est, learned)
