"""Host logic of the learned warm start: seeded training patches, the model
card format and the gate that decides between the learned path and plain
RL (no GPU)."""
import numpy as np
import pytest

from clarity_amd import learned as L
from clarity_amd import synth


def test_training_patch_is_seeded():
    a = L.training_patch(3, convolve=synth.np_convolve)
    b = L.training_patch(3, convolve=synth.np_convolve)
    c = L.training_patch(4, convolve=synth.np_convolve)
    assert a.obs.shape == L.PATCH and a.obs.dtype == np.float32
    assert np.array_equal(a.obs, b.obs) and np.array_equal(a.psf, b.psf)
    assert not np.array_equal(a.obs, c.obs)
    assert L.SIGMA_XY[0] <= a.sigma_xy <= L.SIGMA_XY[1]
    assert L.SIGMA_Z[0] <= a.sigma_z <= L.SIGMA_Z[1]
    assert (a.side_lobe, c.side_lobe) == (0.1, 0.0)
    assert abs(float(a.psf.astype(np.float64).sum()) - 1.0) < 1e-5
    # Poisson counts around the background of 10 plus blurred beads
    assert np.all(a.obs == np.round(a.obs)) and 5 < a.obs.mean() < 40


def test_patch_kernel_is_symmetric():
    """The flips and the y<->x transpose used as augmentation are symmetries
    of the patch kernels (centred at n//2, equal widths in y and x)."""
    obj, psf, *_ = L.patch_object(7)
    z, y, x = psf.shape
    core = psf[1:, 1:, 1:]        # index 0 has no mirror about n//2
    assert np.allclose(core, core[::-1, :, :], atol=1e-9)
    assert np.allclose(core, core[:, ::-1, :], atol=1e-9)
    assert np.allclose(core, core[:, :, ::-1], atol=1e-9)
    assert np.allclose(psf, psf.transpose(0, 2, 1), atol=1e-9)


def _card(err, gap):
    return {"sigma_xy": [1.2, 2.5], "sigma_z": [2.5, 5.0], "yardstick_iterations": 200,
            "refine_iterations": 20,
            "evaluation": {"learned_rel_l2": err, "learned_idiv_rel_gap": gap}}


def test_gate():
    assert L.accepts(_card(5e-4, 1e-4), 2.0, 4.5)
    assert not L.accepts(_card(2e-3, 1e-4), 2.0, 4.5)           # error above the bar
    assert not L.accepts(_card(5e-4, 5e-3), 2.0, 4.5)           # I-divergence gap
    assert not L.accepts(_card(5e-4, 1e-4), 3.0, 4.5)           # widths out of range
    assert not L.accepts(_card(5e-4, 1e-4), 2.0, 5.5)
    assert not L.accepts({"sigma_xy": [1, 3], "sigma_z": [1, 6]}, 2.0, 4.5)  # never measured
    assert L.accepts(_card(0.1, 0.02), 2.0, 4.5, tolerance=0.15, idiv_tolerance=0.05)


def test_card_round_trip(tmp_path):
    p = np.arange(10, dtype=np.float32)
    card = _card(0.1, 0.01)
    card["training"] = {"val_rel_l2": [[0, 1.0], [100, 0.5]]}
    f = tmp_path / "w.npz"
    L.save(str(f), p, card)
    q, c = L.load(str(f))
    assert np.array_equal(p, q) and c == card
    # the file holds no pickles
    with np.load(f, allow_pickle=False) as z:
        assert set(z.files) == {"params", "card"}


def test_committed_weights_match_network():
    """The committed weights have the network's parameter count and a card
    from the yardstick (plain RL) on the 512x512x128 benchmark volume."""
    import os
    if not os.path.exists(L.WEIGHTS):
        pytest.skip("no trained weights committed")
    from oracle import unet as U
    params, card = L.load()
    assert params.size == U.layout()[1] and np.all(np.isfinite(params))
    ev = card["evaluation"]
    assert ev["volume"] == [128, 512, 512] and ev["yardstick_iterations"] == 200
    assert ev["accepted"] == (ev["learned_rel_l2"] <= L.ACCEPT_REL_L2 and
                              ev["learned_idiv_rel_gap"] <= L.ACCEPT_IDIV)
