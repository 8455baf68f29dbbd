"""The fused, pruned power-of-two Richardson-Lucy engine ("fast").

Parity against the CPU oracle (float64 FFT, fp64 reductions) at the small
work shapes, against the generic engine at the headline shape, and the
size-independent invariants of RL (energy, non-negativity, scale
invariance) at 512x512x128.  Tolerances are the RL targets: rel L2 <= 1e-5
after one iteration, <= 1e-4 after up to 100.
"""
import os

import numpy as np
import pytest

import oracle
from clarity_amd import api as capi
from clarity_amd import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(autouse=True)
def _fp64_oracle():
    oracle.set_reduce_mode(oracle.REDUCE_FP64)
    oracle.set_fft_backend(oracle.FFT_OWN)
    yield


def run_plan(api, obs, psf, iters, generic=False, initial=None):
    if generic:
        os.environ["CLARITY_FORCE_GENERIC"] = "1"
    try:
        plan = api.RLPlan(obs.shape, psf.shape)
    finally:
        os.environ.pop("CLARITY_FORCE_GENERIC", None)
    engine = plan.engine
    plan.set_inputs(obs, psf)
    if initial is not None:
        plan.reset(initial)
    plan.iterate(iters)
    out = plan.estimate()
    plan.close()
    return out, engine


def _case(ishape, kshape, seed):
    rng = np.random.default_rng(seed)
    obj = synth.beads_and_filaments(ishape, max(10, int(np.prod(ishape) // 800)), 3, rng)
    psf = synth.gaussian_kernel(kshape, 1.7, 3.2, side_lobe=0.1)
    obs = synth.poisson(synth.np_convolve(obj, psf), rng)
    return obs, psf


SHAPES = [
    ((32, 64, 64), (32, 64, 64)),      # the 64x64x32 beads workload
    ((30, 50, 61), (34, 78, 67)),      # odd image dims, work 64 x 128 x 128
    ((64, 128, 128), (64, 128, 128)),  # a training patch of the warm-start network
    ((48, 96, 96), (48, 96, 96)),      # work 96 x 192 x 192: radix 3
    ((32, 192, 64), (32, 192, 64)),    # work 64 x 384 x 128: radix 6 in y
    ((40, 70, 150), (56, 122, 42)),    # work 96 x 192 x 192 from odd dims
    # kernel deeper than the image: the kernel's planes z < 78 and z >= 178
    # of 256, general z kernel with 16 elements per thread
    ((100, 64, 64), (156, 64, 64)),    # work 256 x 128 x 128
    ((150, 64, 64), (362, 64, 64)),    # work 512 x 128 x 128
]


@pytest.mark.parametrize("ishape,kshape", SHAPES)
def test_fast_vs_oracle(clarity_gpu, ishape, kshape):
    api = clarity_gpu
    obs, psf = _case(ishape, kshape, 21)
    print(f"[fast] image {ishape} kernel {kshape}", flush=True)
    for iters, tol in ((1, 1e-5), (8, 1e-4)):
        got, engine = run_plan(api, obs, psf, iters)
        assert engine == "fast"
        ref = oracle.ml(obs, psf, iters)
        err = rel_l2(got, ref)
        print(f"[fast] iters {iters}: rel L2 {err:.3e}", flush=True)
        assert err <= tol


def _axis_cases():
    """(image, kernel, expected engine work shape): one axis at a time is
    sized so that max(k, i + k/2) lands just below each of the engine's
    lengths along it (the other two axes minimal), so every x, y and z
    kernel instantiation of the default launch shapes runs once."""
    small_i, small_k = {0: (8, 4), 1: (16, 8), 2: (16, 8)}, None
    lengths = {2: [(64, 64, 128), (120, 40, 192), (200, 40, 256), (300, 60, 384),
                   (440, 60, 512), (600, 100, 768), (900, 60, 1024)],
               1: [(64, 64, 128), (120, 40, 192), (200, 40, 256), (300, 60, 384),
                   (440, 60, 512), (600, 100, 768), (900, 60, 1024), (1500, 100, 2048)],
               0: [(20, 10, 64), (60, 30, 96), (80, 40, 128), (120, 40, 192),
                   (160, 80, 256), (300, 100, 512)]}
    minimal = {0: 64, 1: 128, 2: 128}
    out = []
    for ax, rows in lengths.items():
        for i, k, n in rows:
            ishape = [small_i[a][0] for a in range(3)]
            kshape = [small_i[a][1] for a in range(3)]
            ishape[ax], kshape[ax] = i, k
            work = [minimal[a] for a in range(3)]
            work[ax] = n
            out.append((tuple(ishape), tuple(kshape), tuple(work)))
    return out


@pytest.mark.parametrize("ishape,kshape,work", _axis_cases())
def test_fast_every_axis_length(clarity_gpu, ishape, kshape, work):
    api = clarity_gpu
    obs, psf = _case(ishape, kshape, 31)
    plan = api.RLPlan(obs.shape, psf.shape)
    assert (plan.engine, plan.work_shape) == ("fast", work)
    plan.set_inputs(obs, psf)
    plan.iterate(3)
    got = plan.estimate()
    plan.close()
    err = rel_l2(got, oracle.ml(obs, psf, 3))
    print(f"[fast] image {ishape} kernel {kshape} on {work}: rel L2 {err:.3e}", flush=True)
    assert err <= 1e-4


def test_fast_selected_only_for_supported_shapes(clarity_gpu):
    """The fused engine transforms over the smallest of its sizes that is
    at least max(k, i + k/2) per axis, so shapes whose i + k is no size of
    it run on it too; beyond its largest sizes the generic engine runs.
    With CLARITY_REFERENCE_DOMAIN it keeps i + k."""
    api = clarity_gpu
    for ishape, kshape, engine, work in (
            ((32, 64, 64), (32, 64, 64), "fast", (64, 128, 128)),
            ((32, 64, 64), (31, 64, 64), "fast", (64, 128, 128)),   # i + k = 63 in z
            ((16, 100, 37), (16, 28, 91), "fast", (64, 128, 128)),
            ((128, 512, 512), (128, 512, 512), "fast", (256, 768, 768)),
            ((16, 64, 1100), (4, 4, 4), "generic", (20, 68, 1104))):  # x: 1102 > 1024
        p = api.RLPlan(ishape, kshape)
        assert (p.engine, p.work_shape) == (engine, work), (ishape, kshape)
        p.close()
    os.environ["CLARITY_REFERENCE_DOMAIN"] = "1"
    try:
        for kz, engine in ((32, "fast"), (31, "generic")):
            p = api.RLPlan((32, 64, 64), (kz, 64, 64))
            assert p.engine == engine, kz
            p.close()
    finally:
        os.environ.pop("CLARITY_REFERENCE_DOMAIN", None)


@pytest.mark.parametrize("ishape,kshape", [
    ((32, 64, 64), (31, 64, 64)),    # reference domain 63 x 128 x 128
    ((16, 100, 37), (16, 28, 91)),   # reference domain 32 x 128 x 128
    ((1, 128, 100), (1, 9, 7)),      # one plane: 2 x 137 x 107 -> 64 x 256 x 128
    ((5, 40, 30), (3, 120, 90)),     # kernel wider than the image in y and x
    ((7, 33, 45), (20, 17, 9)),      # kernel deeper than the image, odd sizes
])
def test_fast_on_shapes_outside_its_sizes(clarity_gpu, ishape, kshape):
    """Shapes whose reference work domain (i + k) the fused engine has no
    kernels for run on it over a larger alias-free domain, with the
    reference's results (the oracle transforms over i + k)."""
    api = clarity_gpu
    obs, psf = _case(ishape, kshape, 8)
    got, engine = run_plan(api, obs, psf, 6)
    assert engine == "fast"
    err = rel_l2(got, oracle.ml(obs, psf, 6))
    print(f"[fast] image {ishape} kernel {kshape}, 6 it: rel L2 {err:.3e}", flush=True)
    assert err <= 1e-4


def test_fast_matches_generic_100_iterations_beads64(clarity_gpu):
    api = clarity_gpu
    case = synth.make_workload("beads-64x64x32")
    a, ea = run_plan(api, case.obs, case.psf, 100)
    b, eb = run_plan(api, case.obs, case.psf, 100, generic=True)
    assert (ea, eb) == ("fast", "generic")
    ref = oracle.ml(case.obs, case.psf, 100)
    print(f"[fast] beads 64x64x32 100 it: fast {rel_l2(a, ref):.3e} generic {rel_l2(b, ref):.3e}")
    assert rel_l2(a, ref) <= 1e-4
    assert rel_l2(b, ref) <= 1e-4


def test_fast_initial_estimate(clarity_gpu):
    api = clarity_gpu
    case = synth.make_workload("beads-64x64x32")
    init = np.full(case.obs.shape, float(case.obs.mean()), np.float32)
    got, engine = run_plan(api, case.obs, case.psf, 5, initial=init)
    assert engine == "fast"
    assert rel_l2(got, oracle.ml_ex(case.obs, case.psf, 5, initial=init)) <= 1e-4


def test_fast_headline_shape_invariants(clarity_gpu):
    """512x512x128 image, work 1024x1024x256: fast vs generic engine after
    3 iterations, energy, non-negativity and scale invariance."""
    api = clarity_gpu
    shape = (128, 512, 512)
    rng = np.random.default_rng(3)
    obj = synth.beads_and_filaments(shape, 16000, 2400, rng)
    psf = synth.gaussian_kernel(shape, 2.0, 4.5, side_lobe=0.1)
    obs = synth.poisson(api.Convolve(obj, psf), rng)
    a, ea = run_plan(api, obs, psf, 3)
    b, eb = run_plan(api, obs, psf, 3, generic=True)
    assert (ea, eb) == ("fast", "generic")
    err = rel_l2(a, b)
    print(f"[fast] 512x512x128, 3 it: fast vs generic rel L2 {err:.3e}", flush=True)
    assert err <= 1e-5
    assert np.all(a >= 0)
    assert abs(a.astype(np.float64).sum() / obs.astype(np.float64).sum() - 1) < 1e-5
    c, _ = run_plan(api, 8 * obs, psf, 3)
    assert rel_l2(c, 8 * a) <= 1e-5


def test_fast_timing_api(clarity_gpu):
    api = clarity_gpu
    case = synth.make_workload("beads-64x64x32")
    plan = api.RLPlan(case.obs.shape, case.psf.shape)
    plan.set_inputs(case.obs, case.psf)
    kinds = plan.describe()
    assert [k[0] for k in kinds] == ["x_fwd", "y_fwd", "z_mod", "y_inv", "x_mid",
                                     "x_last", "finish", "x_last_fwd"]
    assert [k[2] for k in kinds] == [0, 2, 2, 2, 1, 0, 1, 1]
    plan.set_timing(True)
    plan.timings(reset=True)
    plan.iterate(3)
    plan.iterate(1)
    t = plan.timings()
    # x_fwd and x_last once per call; x_last_fwd joins iterations of a call
    assert [c for c, _ in t] == [2, 8, 8, 8, 4, 2, 4, 2]
    assert all(ms > 0 for _, ms in t)
    plan.close()


@pytest.mark.parametrize("shape", [(128, 512, 512), (32, 64, 64)])
def test_x_tile_order_leaves_results_unchanged(clarity_gpu, shape):
    """The x passes take their tiles from the last plane down by default
    (xr=1) and from the first up with xr=0: each tile computes the same
    values and writes its partial sum to its own slot, so the RL result is
    bit for bit the same (512x512x128: 16-row tiles; 64x64x32: 8-row tiles
    paired per 128-byte span)."""
    api = clarity_gpu
    rng = np.random.default_rng(9)
    obs = rng.uniform(1.0, 100.0, shape).astype(np.float32)
    psf = synth.gaussian_kernel(shape, 2.0, 4.5, side_lobe=0.1)
    outs = []
    for tune in ("xr=0", "xr=1"):
        os.environ["CLARITY_FAST_TUNE"] = tune
        try:
            outs.append(api.MaximumLikelihoodDeconvolve(obs, psf, 3))
            assert api.LastEngine() == "fast"
        finally:
            os.environ.pop("CLARITY_FAST_TUNE", None)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("tune,exact", [("xh=0", True), (None, False)])
def test_fused_x_boundary_matches_split(clarity_gpu, tune, exact):
    """Iterations chained through x_last_fwd equal the same iterations run
    one call at a time (x_last, then x_fwd).  Without the same-size x
    specialisations (xh=0) both routes run the same loads and arithmetic, so
    bit for bit; with them (the default) x_fwd and x_last may contract the
    butterflies of the known-zero half differently, so to float rounding."""
    api = clarity_gpu
    case = synth.make_workload("beads-64x64x32")
    if tune is not None:
        os.environ["CLARITY_FAST_TUNE"] = tune
    try:
        plan = api.RLPlan(case.obs.shape, case.psf.shape)
    finally:
        os.environ.pop("CLARITY_FAST_TUNE", None)
    plan.set_inputs(case.obs, case.psf)
    plan.reset()
    plan.iterate(5)
    chained = plan.estimate()
    plan.reset()
    for _ in range(5):
        plan.iterate(1)
    split = plan.estimate()
    plan.close()
    if exact:
        assert np.array_equal(chained, split)
    else:
        assert rel_l2(chained, split) <= 1e-6


@pytest.mark.parametrize("ishape,tune", [
    ((32, 64, 64), "zw=3"),                # z 64, 8 per thread; 520 tiles over 174 groups
    ((64, 96, 96), "zw=2,ze=16,zt=16"),    # z 128, 16 per thread
    ((48, 64, 64), "zw=4"),                # z 96, radix 3
    ((96, 48, 64), "zw=2,zt=32"),          # z 192, radix 3
    ((128, 64, 64), "zw=4"),               # z 256, 16 per thread, 32-column tiles
    ((128, 64, 64), "zw=5,ze=8,zt=32"),    # z 256, 8 per thread
    ((128, 64, 64), "zw=8,ze=16,zt=16"),   # z 256, 16-column tiles
])
def test_z_pipelined_matches_one_tile(clarity_gpu, ishape, tune):
    """The same-size z pass over several column tiles per workgroup (the next
    tile's planes loaded during the current tile's inverse transform, tile
    counts that do not divide evenly) gives the one-tile kernel's result to
    float rounding, and the oracle's."""
    api = clarity_gpu
    obs, psf = _case(ishape, ishape, 17)
    outs = {}
    for t in ("zw=1", tune):
        os.environ["CLARITY_FAST_TUNE"] = t
        try:
            plan = api.RLPlan(obs.shape, psf.shape)
        finally:
            os.environ.pop("CLARITY_FAST_TUNE", None)
        assert plan.engine == "fast" and plan.work_shape[0] == 2 * ishape[0]
        plan.set_inputs(obs, psf)
        plan.reset()
        plan.iterate(5)
        outs[t] = plan.estimate()
        plan.close()
    assert rel_l2(outs[tune], outs["zw=1"]) <= 1e-6
    assert rel_l2(outs[tune], oracle.ml(obs, psf, 5)) <= 1e-4


@pytest.mark.parametrize("path", ["grouped", "streams"])
def test_batch_uneven_volumes(clarity_gpu, path):
    """Batched RL over uneven volumes equals per-volume RL.  Grouped (the
    default when every volume is on the fused engine): one launch per pass
    and kernel instantiation over all volumes, each volume running its own
    engine's kernels, arguments and order, so bit for bit the single-volume
    call.  Streams (CLARITY_BATCH_STREAMS): each volume on its own plan."""
    api = clarity_gpu
    rng = np.random.default_rng(44)
    shapes = [(32, 64, 64), (16, 48, 40), (24, 96, 64), (32, 64, 64), (8, 20, 30),
              (48, 128, 96), (8, 20, 30)]
    imgs, ks = [], []
    for i, s in enumerate(shapes):
        obj = synth.beads_and_filaments(s, 30, 2, rng)
        k = synth.gaussian_kernel(s, rng.uniform(1.2, 2.5), rng.uniform(2.5, 5.0))
        imgs.append(synth.poisson(synth.np_convolve(obj, k), rng))
        ks.append(k)
    if path == "streams":
        os.environ["CLARITY_BATCH_STREAMS"] = "1"
    try:
        outs = api.MaximumLikelihoodDeconvolveBatch(imgs, ks, 6)
    finally:
        os.environ.pop("CLARITY_BATCH_STREAMS", None)
    for a, k, o in zip(imgs, ks, outs):
        single = api.MaximumLikelihoodDeconvolve(a, k, 6)
        if path == "grouped":
            assert np.array_equal(o, single)
        else:
            assert rel_l2(o, single) <= 1e-6
        assert rel_l2(o, oracle.ml(a, k, 6)) <= 1e-4


def test_graph_replay_matches_eager(clarity_gpu):
    """An iteration captured as a HIP graph and replayed gives bit-identical
    results to eager launches, replay after replay, with eager work (the
    estimate copy-out and a host sync) between the replays; iteration counts
    below, at and past the graphs of 8 chained iterations (3, 9 = 8 + last,
    17 = 2 x 8 + last, 8 = 7 single + last)."""
    api = clarity_gpu
    case = synth.make_workload("beads-64x64x32")
    os.environ["CLARITY_NO_GRAPH"] = "1"
    try:
        eager = api.RLPlan(case.obs.shape, case.psf.shape)
    finally:
        os.environ.pop("CLARITY_NO_GRAPH", None)
    graph = api.RLPlan(case.obs.shape, case.psf.shape)
    for p in (eager, graph):
        p.set_inputs(case.obs, case.psf)
    for k, n in enumerate((3, 9, 17, 8)):
        eager.iterate(n)
        graph.iterate(n)
        a, b = eager.estimate(), graph.estimate()
        assert np.array_equal(a, b), (k, n)
        assert eager.scale() == graph.scale()
    eager.close()
    graph.close()


def test_autotune_leaves_results_unchanged(clarity_gpu):
    """Autotuning times launch shapes on the plan's own inputs; the estimate
    it found is restored bit for bit, and iterating with the chosen shapes
    gives the same RL result to float rounding."""
    api = clarity_gpu
    # a shape no other test uses: the choice is cached for the process
    obs, psf = _case((24, 64, 64), (40, 64, 64), 5)
    plan = api.RLPlan(obs.shape, psf.shape)
    assert plan.engine == "fast"
    plan.set_inputs(obs, psf)
    plan.reset()
    plan.iterate(4)
    before = plan.estimate()
    choice = plan.autotune(2)
    print(f"[fast] autotune image 64x64x24 kernel 64x64x40: {choice}", flush=True)
    assert "zt=" in choice and "fx=" in choice
    assert np.array_equal(plan.estimate(), before)
    plan.reset()
    plan.iterate(4)
    assert rel_l2(plan.estimate(), before) <= 1e-5
    plan.close()
    # a new plan of the same shape starts from the tuned shapes
    again = api.RLPlan(obs.shape, psf.shape)
    again.set_inputs(obs, psf)
    again.iterate(4)
    assert rel_l2(again.estimate(), before) <= 1e-5
    again.close()


def test_plan_outlives_unregister():
    """A plan holds a registration: after the caller's last
    Clarity_UnRegister the plan still iterates (its twiddle tables and the
    library stream are kept) and gives the result it gives registered;
    destroying it then tears the runtime down, and the next call registers
    again.  Run in a fresh process, so that the session's registration does
    not hide the teardown."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "from clarity_amd import api, synth\n"
        "c = synth.make_workload('beads-64x64x32')\n"
        "api.Register()\n"
        "ref = api.MaximumLikelihoodDeconvolve(c.obs, c.psf, 6)\n"
        "p = api.RLPlan(c.obs.shape, c.psf.shape)\n"
        "p.set_inputs(c.obs, c.psf)\n"
        "p.iterate(3)\n"
        "api.UnRegister()\n"
        "p.iterate(3)\n"
        "out = p.estimate()\n"
        "p.close()\n"
        "e = np.linalg.norm(out.astype(np.float64) - ref) / np.linalg.norm(ref)\n"
        "print('plan after UnRegister: rel', e, flush=True)\n"
        "again = api.MaximumLikelihoodDeconvolve(c.obs, c.psf, 6)\n"
        "print('implicit re-registration: equal', np.array_equal(again, ref), flush=True)\n"
        "sys.exit(0 if e <= 1e-6 and np.array_equal(again, ref) else 3)\n" % root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
