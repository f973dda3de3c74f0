"""The drop-in tile call from many host threads, as the reference's worker
pool calls sp_PathTraceTile (src/main.cpp:731-759: up to 16 threads,
src/config.h:21, pull 64x64 tiles from a shared counter and trace them on one
shared context into one image).

Concurrent spg_path_trace_tile calls with the same context are combined
into shared wavefronts; the image must equal one spg_path_trace_tiles call
over all tiles bit for bit (per-pixel random streams do not depend on which
tiles are traced together), and the callers' metrics must add up to that
call's counters.  Calls with different contexts (here: different image
sizes) are traced apart and each equals its own single-caller render.
"""
import threading

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes

pytestmark = pytest.mark.gpu


def _workers(ctx, tiles, img, threads, repeat=1):
    """`threads` workers pull tile indices from a shared counter and call
    the drop-in tile entry point; returns each worker's metrics."""
    lock = threading.Lock()
    nxt = [0]
    mets = [V.Metrics() for _ in range(threads)]
    errors = []
    total = len(tiles) * repeat

    def work(k):
        try:
            while True:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= total:
                    return
                V.sp_PathTraceTile(ctx, tiles[i % len(tiles)], img, mets[k])
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    return mets


@pytest.mark.parametrize("rng,est", [(V.RNG_PCG, V.EST_PARITY),
                                     (V.RNG_XORSHIFT_PIXEL, V.EST_UNBIASED)])
def test_sixteen_workers_equal_one_batched_call(device, rng, est):
    desc = scenes.config1(256, 256, env="cube")
    s, cam = scenes.load_gpu(device, desc)
    tiles = V.ComputeTiles(256, 256, 64, 64)
    ctx = V.make_context(cam, s, spp=4, bounces=3, clamp=10.0, estimator=est,
                         rng=rng, seed=0x5EED)
    want = np.zeros(256 * 256 * 4, np.float32)
    mw = V.path_trace_tiles(ctx, tiles, want)
    img = np.full(256 * 256 * 4, np.nan, np.float32)
    mets = _workers(ctx, tiles, img, 16)
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))
    for k in (V.METRIC_PATHS, V.METRIC_RAYS, V.METRIC_HITS, V.METRIC_MISSES):
        assert sum(m.values[k] for m in mets) == mw.values[k], k
    assert all(m.values[0] > 0 for m in mets if m.values[V.METRIC_PATHS])
    s.close()


def test_two_contexts_of_different_sizes_from_threads(device):
    """Two contexts on one device whose images differ in size, traced from
    interleaved threads: no call may see the other's staging buffer."""
    big = scenes.config1(256, 192, env="cube")
    small = scenes.config1(96, 64, env="cube")
    sb, cb = scenes.load_gpu(device, big)
    ss, cs = scenes.load_gpu(device, small)
    cb_ctx = V.make_context(cb, sb, spp=2, bounces=3, estimator=V.EST_PARITY,
                            rng=V.RNG_PCG, seed=11)
    cs_ctx = V.make_context(cs, ss, spp=2, bounces=3, estimator=V.EST_PARITY,
                            rng=V.RNG_PCG, seed=12)
    tb = V.ComputeTiles(256, 192, 32, 32)
    tsm = V.ComputeTiles(96, 64, 32, 32)
    want_b = np.zeros(256 * 192 * 4, np.float32)
    want_s = np.zeros(96 * 64 * 4, np.float32)
    V.path_trace_tiles(cb_ctx, tb, want_b)
    V.path_trace_tiles(cs_ctx, tsm, want_s)
    img_b = np.zeros_like(want_b)
    img_s = np.zeros_like(want_s)
    errors = []

    def run(ctx, tiles, img, repeat):
        try:
            for _ in range(repeat):
                for t in tiles:
                    V.sp_PathTraceTile(ctx, t, img)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=run, args=(cb_ctx, tb, img_b, 2)) for _ in range(4)]
    ts += [threading.Thread(target=run, args=(cs_ctx, tsm, img_s, 6)) for _ in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert np.array_equal(img_b.view(np.uint32), want_b.view(np.uint32))
    assert np.array_equal(img_s.view(np.uint32), want_s.view(np.uint32))
    sb.close()
    ss.close()


def test_worker_errors_reach_every_caller(device):
    """A context the tracer rejects (bounces 0) fails in every worker's own
    thread with the library's message."""
    desc = scenes.config1(64, 64, env="cube")
    s, cam = scenes.load_gpu(device, desc)
    ctx = V.make_context(cam, s, spp=1, bounces=3, estimator=V.EST_PARITY,
                         rng=V.RNG_PCG, seed=1)
    ctx.config.bounces = 0
    img = np.zeros(64 * 64 * 4, np.float32)
    msgs = []

    def run():
        try:
            V.sp_PathTraceTile(ctx, V.Tile(0, 0, 32, 32), img)
        except V.SpgError as e:
            msgs.append(str(e))

    ts = [threading.Thread(target=run) for _ in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len(msgs) == 8 and all("bounces" in m for m in msgs), msgs
    s.close()
