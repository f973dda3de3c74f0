"""The learned guide on the GPU: the MFMA network against the FP64 oracle
and an FP16-emulating forward pass, the training kernels' gradient and Adam
step against the oracle, learning on a known target, and the guided
estimator: unbiased (closed forms, agreement with the plain tracer) and
lower variance once trained (what the network gains over BSDF sampling)."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _guide import (fp16_forward, oracle_forward, oracle_loss_grad, random_inputs,
                    random_params, random_samples, unit, W_OFF, ROWS)
from _libs import GUIDE_PARAMS, SpoLobes, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture
def guide(device):
    g = V.Guide(device, seed=3, max_weight=0.0)
    yield g
    g.close()


def test_forward_matches_fp16_emulation_and_oracle(guide):
    p = random_params(1, scale_out=2.0)
    guide.set_params(p)
    assert np.array_equal(guide.params(), p)
    x = random_inputs(2, 1000)   # 15 full wave tiles + a partial one
    out, lobes = guide.evaluate(x)
    emu = fp16_forward(p, x)
    # layout and rounding points: FP32 accumulation order is the only
    # difference left, which can move a hidden activation by one FP16 ulp
    err = np.abs(out - emu)
    assert np.max(err / (np.abs(emu) + 1.0)) < 1e-2
    assert np.mean(err) < 1e-3
    # against exact arithmetic: within FP16 rounding of 64-wide layers
    ref = oracle_forward(p, x)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 1e-2
    # the mixture read from the outputs
    O = oracle()
    for j in range(0, 1000, 97):
        L = SpoLobes()
        O.spo_guide_lobes(out[j].astype(np.float64), L)
        assert lobes[j, 0:8] == pytest.approx(list(L.pi), rel=1e-4, abs=1e-6)
        assert lobes[j, 8:16] == pytest.approx(list(L.kappa), rel=1e-4)
        mu = np.array([list(m) for m in L.mu]).ravel()
        assert lobes[j, 16:40] == pytest.approx(mu, abs=1e-5)


def _layer_views(g):
    for l in range(5):
        o = W_OFF[l]
        yield g[o:o + ROWS[l] * 65]


def test_gradient_matches_oracle(guide):
    p = random_params(4, scale_out=1.0)
    guide.set_params(p)
    s = random_samples(5, 700)
    g, loss = guide.gradient(s)
    lref, gref = oracle_loss_grad(p, s)
    assert loss == pytest.approx(lref, rel=2e-2)
    cos = g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref))
    assert cos > 0.999
    assert np.linalg.norm(g - gref) / np.linalg.norm(gref) < 3e-2
    for gl, rl in zip(_layer_views(g), _layer_views(gref)):
        assert np.linalg.norm(gl - rl) / np.linalg.norm(rl) < 6e-2


def test_gradient_normalization_and_clamp(guide):
    p = random_params(6)
    guide.set_params(p)
    s = random_samples(7, 256)
    w = s[:, 12].astype(np.float64)
    wn = np.minimum(w * len(w) / w.sum(), 2.0)
    g, _ = guide.gradient(s, normalize=True, max_weight=2.0)
    _, gref = oracle_loss_grad(p, s, weights=wn)
    assert np.linalg.norm(g - gref) / np.linalg.norm(gref) < 3e-2


def test_adam_step_matches_oracle(device):
    """One training step = the batch gradient (pinned above) followed by
    Adam: the parameters equal the oracle's Adam step applied to the
    gradient the GPU computes for the same batch."""
    lr = 1e-3
    g = V.Guide(device, seed=1, learning_rate=lr, batch_size=1024, max_weight=0.0)
    p = random_params(8)
    g.set_params(p)
    s = random_samples(9, 1024)
    grad, _ = g.gradient(s, normalize=True)
    g.train(s, steps=1)
    p1 = g.params()
    q = p.copy()
    m = np.zeros(GUIDE_PARAMS, np.float32)
    v = np.zeros(GUIDE_PARAMS, np.float32)
    oracle().spo_guide_adam(q, m, v, grad.astype(np.float64), GUIDE_PARAMS, lr, 0.9,
                            0.99, 1e-8, 1)
    # the first step moves every parameter by lr * g / (|g| + eps); where g
    # is at the level of the FP32 atomics' summation-order noise the two
    # gradient evaluations may differ in sign
    sel = np.abs(grad) > 1e-5 * np.abs(grad).max()
    assert sel.mean() > 0.9
    assert np.max(np.abs(p1 - q)[sel]) < 2e-6
    assert np.max(np.abs(p1 - q)) <= 2 * lr * 1.001
    moved = np.abs(p1 - p)
    assert np.median(moved) == pytest.approx(lr, rel=0.05)
    # and against the oracle's own gradient: the same direction
    w = s[:, 12].astype(np.float64)
    _, gref = oracle_loss_grad(p, s, weights=w * len(w) / w.sum())
    big = np.abs(gref) > 0.05 * np.abs(gref).max()
    assert np.all(np.sign(p - p1)[big] == np.sign(gref)[big])
    g.close()


def test_training_learns_a_known_distribution(device):
    """Directions drawn from vMF(mu(pn), 50) with mu pointing from the hit
    to a fixed point: the loss falls from the uniform entropy log(4 pi) to
    near the target's entropy, and the learned density peaks there."""
    g = V.Guide(device, seed=2, learning_rate=5e-3, batch_size=4096, max_weight=0.0)
    rng = np.random.default_rng(0)
    n = 32768
    s = np.zeros((n, 13), np.float32)
    s[:, 0:9] = random_inputs(11, n)
    target = np.array([0.5, 1.5, 0.5])
    mu = unit(target - s[:, 0:3])
    O = oracle()
    for j in range(n):
        w = np.zeros(3)
        O.spo_vmf_sample(50.0, np.ascontiguousarray(mu[j]), rng.uniform(),
                         rng.uniform(), w)
        s[j, 9:12] = w
    s[:, 12] = 1.0
    first = g.train(s, steps=1)
    last = g.train(s, steps=400)
    assert first > 1.5
    assert last < -0.3  # target entropy: 1 - log(50/2pi) = -1.07
    x = random_inputs(12, 256)
    _, lobes = g.evaluate(x)
    mu_t = unit(target - x[:, 0:3])
    dens = []
    for j in range(256):
        L = lobes[j]
        q = 0.0
        for k in range(8):
            kap = L[8 + k]
            d = L[16 + 3 * k:19 + 3 * k] @ mu_t[j]
            q += L[k] * kap / (2 * np.pi * -np.expm1(-2 * kap)) * np.exp(kap * (d - 1))
        dens.append(q)
    assert np.median(dens) > 2.0  # uniform would be 0.08


def test_training_steps_are_deterministic(device):
    """Fixed-order reductions (weight sum, per-chunk partial gradients
    added in chunk order): the same batches give bit-identical gradients
    and parameters on two guides and on repeated calls."""
    s = random_samples(21, 6000)
    a = V.Guide(device, seed=5, batch_size=2048, max_weight=10.0)
    b = V.Guide(device, seed=5, batch_size=2048, max_weight=10.0)
    g1, _ = a.gradient(s, normalize=True, max_weight=10.0)
    g2, _ = a.gradient(s, normalize=True, max_weight=10.0)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32))
    a.train(s, steps=5)
    b.train(s, steps=5)
    assert np.array_equal(a.params().view(np.uint32), b.params().view(np.uint32))
    a.close()
    b.close()


def test_checkpoint_resumes_training_exactly(device, tmp_path):
    """spg_guide_save / spg_guide_load carry the parameters, both Adam
    moments, the step count and the configuration: a guide restored from a
    checkpoint continues training bit-identically to the one that wrote it,
    while one restored from parameters alone (spg_guide_set_params, Adam
    state reset) does not."""
    s = random_samples(22, 8192)
    a = V.Guide(device, seed=7, learning_rate=2e-3, batch_size=2048, max_weight=10.0)
    a.train(s, steps=6)
    path = tmp_path / "guide.ckpt"
    a.save(path)
    b = V.Guide(device, seed=99, learning_rate=9e-3, batch_size=1024, max_weight=0.0)
    b.load(path)
    assert np.array_equal(a.params().view(np.uint32), b.params().view(np.uint32))
    assert b.stats()["steps"] == a.stats()["steps"] == 6
    c = V.Guide(device, seed=7, learning_rate=2e-3, batch_size=2048, max_weight=10.0)
    c.set_params(a.params())
    a.train(s, steps=5)
    b.train(s, steps=5)
    c.train(s, steps=5)
    pa, pb, pc = a.params(), b.params(), c.params()
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert not np.array_equal(pa, pc)
    # a file that is not a checkpoint is rejected and leaves the guide as it was
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"SPGGUIDE" + bytes(40))
    with pytest.raises(V.SpgError):
        b.load(bad)
    assert np.array_equal(b.params().view(np.uint32), pb.view(np.uint32))
    for g in (a, b, c):
        g.close()


_loaded = {}


def _render(device, desc, spp, est, guide=None, train=False, offset=0, seed=0x5EED):
    key = (id(device), id(desc))
    if key not in _loaded:
        _loaded.clear()
        _loaded[key] = scenes.load_gpu(device, desc) + (desc,)
    s, cam, _ = _loaded[key]
    w, h = desc["camera"][0], desc["camera"][1]
    ctx = V.make_context(cam, s, spp=spp, bounces=6, clamp=0.0, estimator=est,
                         rng=V.RNG_PCG, seed=seed, sample_offset=offset,
                         transcendentals=V.TX_FAST, guide=guide, train=train)
    img = np.zeros(w * h * 4, np.float32)
    V.path_trace_tiles(ctx, V.ComputeTiles(w, h, 64, 64), img)
    return img.reshape(h, w, 4)[:, :, :3]


@pytest.fixture(autouse=True)
def _drop_scenes():
    yield
    for s, _, _ in _loaded.values():
        s.close()
    _loaded.clear()


def test_guided_needs_a_guide(device):
    desc = scenes.config1(16, 16, env="cube", env_size=16)
    with pytest.raises(V.SpgError):
        _render(device, desc, 1, V.EST_GUIDED)


def test_guided_furnace_closed_form(device):
    """Inside a closed sphere of albedo a and emission E, every estimator
    has mean E (1 - q^N) / (1 - q) up to the Fresnel term; the guided
    one with an untrained network included."""
    v, i = V.gen_icosphere(3)
    vin = v.copy()
    vin[:, 3:6] *= -1
    iin = i.reshape(-1, 3)[:, ::-1].ravel().copy()
    a, E, N = 0.5, 1.0, 6
    desc = {"meshes": [(vin, iin, True)],
            "objects": [(0, 1, (0, 0, 0), (0, 0, 0, 1), (5, 5, 5))],
            "materials": [(1, (a, a, a), (E, E, E), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE)],
            "background": 9, "camera": (32, 32, (0, 0, 0), (0, 0, 0, 1), 0.8)}
    g = V.Guide(device, seed=5)
    gm = _render(device, desc, 64, V.EST_GUIDED, guide=g)[:, :, 0].mean()
    um = _render(device, desc, 64, V.EST_UNBIASED)[:, :, 0].mean()
    q = a * (1 - 0.04)
    hi = E * (1 - q ** N) / (1 - q)
    assert E < gm < hi and gm == pytest.approx(hi, rel=0.05)
    assert gm == pytest.approx(um, rel=0.01)
    g.close()


def test_guided_is_unbiased_against_plain(device):
    """Sky + sun over the icosphere scene: guided (untrained and trained)
    and plain images have equal means within their Monte Carlo error."""
    desc = scenes.config1(64, 64, env="cube", env_size=32)
    plain = [_render(device, desc, 64, V.EST_UNBIASED, seed=k).mean() for k in range(4)]
    g = V.Guide(device, seed=7, batch_size=4096)
    untrained = [_render(device, desc, 64, V.EST_GUIDED, guide=g, seed=10 + k).mean()
                 for k in range(4)]
    for k in range(20):
        _render(device, desc, 1, V.EST_GUIDED, guide=g, train=True, offset=1000 + k)
    assert g.stats()["steps"] > 0
    trained = [_render(device, desc, 64, V.EST_GUIDED, guide=g, seed=20 + k).mean()
               for k in range(4)]
    pm = np.mean(plain)
    sd = np.std(plain + untrained + trained) + 1e-3 * pm
    assert abs(np.mean(untrained) - pm) < 4 * sd
    assert abs(np.mean(trained) - pm) < 4 * sd
    g.close()


def _lamp_scene(w, h):
    """A 20x20 diffuse floor lit only by a small emissive sphere (radius
    0.25, radiance 20) at height 2: cosine sampling finds the lamp with
    probability ~1% per bounce, a guide that points at it almost always."""
    pv, pidx = V.gen_plane()
    sv, sidx = V.gen_icosphere(3)
    qfloor = scenes.quat_axis_angle((1, 0, 0), -np.pi / 2)
    return {"meshes": [(pv, pidx, False), (sv, sidx, True)],
            "objects": [(0, 1, (0, 0, 0), tuple(qfloor), (10, 10, 10)),
                        (1, 2, (0, 2, 0), (0, 0, 0, 1), (0.25, 0.25, 0.25))],
            "materials": [(1, (0.5, 0.5, 0.5), (0, 0, 0), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE),
                          (2, (0, 0, 0), (20, 20, 20), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE),
                          (9, (0, 0, 0), (0, 0, 0), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE)],
            "background": 9,
            "camera": (w, h, (0, 3, 5), tuple(scenes.quat_axis_angle((1, 0, 0), -0.5)),
                       0.8)}


def test_trained_guide_reduces_variance(device):
    """What the trained network gains over the method without it: at equal
    samples the per-pixel variance of the guided estimator after online
    training is far below cosine (BSDF) sampling's and below the untrained
    guide's, at the same mean; the learned lobe at a floor point points at
    the lamp."""
    desc = _lamp_scene(128, 96)
    reps, spp = 8, 4

    def stats(est, guide=None, base=0):
        imgs = np.stack([_render(device, desc, spp, est, guide=guide,
                                 seed=base + k).mean(axis=2) for k in range(reps)])
        return imgs.mean(), imgs.var(axis=0, ddof=1).mean()

    m_plain, v_plain = stats(V.EST_UNBIASED, base=100)
    g = V.Guide(device, seed=9, batch_size=16384, steps_per_pass=4)
    m_untr, v_untr = stats(V.EST_GUIDED, guide=g, base=200)
    for k in range(60):
        _render(device, desc, 1, V.EST_GUIDED, guide=g, train=True, offset=5000 + k)
    assert g.stats()["collected"] > 1000
    m_tr, v_tr = stats(V.EST_GUIDED, guide=g, base=300)
    print(f"\nmeans plain {m_plain:.5f} untrained {m_untr:.5f} trained {m_tr:.5f}; "
          f"variance plain {v_plain:.3e} untrained {v_untr:.3e} trained {v_tr:.3e}; "
          f"gain {v_plain / v_tr:.2f}x")
    assert v_plain / v_tr > 4.0
    assert v_untr / v_tr > 4.0
    se = np.sqrt((v_plain + v_tr) / (reps * 128 * 96))
    assert abs(m_tr - m_plain) < 6 * se + 0.01 * m_plain
    # floor point under the lamp: pn from the scene bounds
    lo, hi = np.array([-10.0, 0.0, -10.0]), np.array([10.0, 2.25, 10.0])
    x = np.zeros((1, 9), np.float32)
    x[0, 0:3] = (np.array([0.5, 0.0, 0.5]) - lo) / (hi - lo)
    x[0, 4] = 1.0
    x[0, 7] = 1.0
    _, lobes = g.evaluate(x)
    k = int(np.argmax(lobes[0, 0:8]))
    want = unit(np.array([0.0, 2.0, 0.0]) - np.array([0.5, 0.0, 0.5]))
    assert lobes[0, 16 + 3 * k:19 + 3 * k] @ want > 0.9
    assert lobes[0, 8 + k] > 30.0
    g.close()


def test_online_training_is_reproducible(device):
    """Online training inside trace calls collects its samples in a fixed
    order (a scan over a Weyl permutation of the paths, not atomic appends)
    and reduces gradients in a fixed order: two guides trained by the same
    passes end bit-identical, and so do the images they render."""
    desc = scenes.config1(96, 64, env="cube", env_size=32)
    guides = [V.Guide(device, seed=11, batch_size=4096, steps_per_pass=2)
              for _ in range(2)]
    imgs, collected = [], []
    for g in guides:
        for k in range(6):
            _render(device, desc, 1, V.EST_GUIDED, guide=g, train=True, offset=700 + k)
        collected.append(g.samples())
        imgs.append(_render(device, desc, 4, V.EST_GUIDED, guide=g, seed=77))
    assert guides[0].stats()["steps"] == 12
    assert len(collected[0]) > 1000
    assert np.array_equal(collected[0].view(np.uint32), collected[1].view(np.uint32))
    assert np.array_equal(guides[0].params().view(np.uint32),
                          guides[1].params().view(np.uint32))
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    for g in guides:
        g.close()
