"""Batched rendering: many jobs of uneven resolution, sample count and depth
in one call (shared queues, per-job segments, one captured kernel sequence
replayed per batch) give every job the bit-identical sums it gets alone, and
those match the CPU oracle's statistics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _relmse(x, r):
    ok = np.isfinite(x) & np.isfinite(r)
    return float(np.mean((x[ok] - r[ok]) ** 2 / (r[ok] ** 2 + 1e-2)))


def test_jobs_equal_single_renders(oracle):
    from xenon_amd import GpuScene, HipRenderer, HipScene, scenes
    specs = [scenes.Job(0, 48, 32, 12, 6, 1), scenes.Job(1, 40, 24, 5, 8, 2), scenes.Job(2, 32, 32, 7, 10, 3),
             scenes.Job(0, 16, 48, 3, 4, 4), scenes.Job(1, 64, 16, 9, 12, 5)]
    descs = [scenes.job_scene(j, small=True) for j in specs]
    hs = [HipScene(d) for d in descs]
    # a small path capacity forces many batches with several chunks each
    with HipRenderer(32, 32, 8, 3, spp_per_launch=3) as r:
        gs = [GpuScene(r, h) for h in hs]
        jobs = [dict(scene=g, camera=h.camera(), width=d.width, height=d.height, spp_first=1, spp_count=d.spp,
                     max_bounces=d.max_bounces, min_bounces=d.min_bounces) for g, h, d in zip(gs, hs, descs)]
        print(f"[jobs] {len(jobs)} jobs batched", flush=True)
        _, sums = r.render_jobs(jobs, want_sums=True)
        st = r.stats()
        assert st["paths"] == sum(d.width * d.height * d.spp for d in descs)
        for k, (h, d) in enumerate(zip(hs, descs)):
            print(f"[jobs] single render of job {k}", flush=True)
            with HipRenderer(d.width, d.height, d.max_bounces, d.min_bounces) as one:
                one.upload_scene(h)
                one.render(h.camera(), 1, d.spp, want_image=False)
                alone = one.read_accum()[0]
            assert np.array_equal(alone.view(np.uint32), sums[k].view(np.uint32)), k
            ref, _ = oracle.load(d).render_sum(1, d.spp)
            assert _relmse(sums[k] / d.spp, ref / d.spp) <= 1e-3
        for g in gs:
            g.release()


def test_graph_and_direct_launch_agree():
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(40, 40, spp=10)
    hs = HipScene(desc)
    out = []
    for graph in (True, False):
        with HipRenderer(40, 40, 8, 3, spp_per_launch=2, use_graph=graph) as r:
            r.upload_scene(hs)
            r.render(hs.camera(), 1, 6, want_image=False)
            r.render(hs.camera(), 7, 4, want_image=False)
            out.append(r.read_accum()[0])
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


def test_tuning_does_not_change_results(tmp_path):
    """Workgroup size, LDS stack depth and queue capacity change speed only:
    every sample's path and the per-pixel summation order are fixed, so the
    sums are bit-identical under any tuning, and autotune picks one of the
    candidates and caches it."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(48, 40, spp=12)
    hs = HipScene(desc)
    outs = []
    with HipRenderer(48, 40, 8, 3) as r:
        r.upload_scene(hs)
        for b, l, n in ((256, 16, 0), (64, 8, 3), (128, 32, 5), (192, 1, 12)):
            print(f"[tuning] block {b} lds {l} spp/batch {n}", flush=True)
            r.set_tuning(b, l, n)
            t = r.tuning()
            assert t["block_size"] == b and t["lds_stack"] == l
            r.render(hs.camera(), 1, 12, want_image=False)
            outs.append(r.read_accum()[0])
        for o in outs[1:]:
            assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
        cache = str(tmp_path / "tune.json")
        res = r.autotune(hs.camera(), probe_spp=4, cache=cache)
        assert not res["cached"] and res["candidates"] >= 15 and res["block_size"] in (64, 128, 256)
        again = r.autotune(hs.camera(), probe_spp=4, cache=cache)
        assert again["cached"] and again["lds_stack"] == res["lds_stack"]
        r.render(hs.camera(), 1, 12, want_image=False)
        assert np.array_equal(r.read_accum()[0].view(np.uint32), outs[0].view(np.uint32))


def test_autotune_centres_on_default_batch_and_is_not_slower():
    """The batch-size candidates are half, the default (about 2^28 paths)
    and double; the winner renders the probe no slower than the default
    tuning does, within timing noise."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.big_mesh(width=640, height=360, spp=8, level=5)
    hs = HipScene(desc)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(hs)
        default = r.tuning()
        print(f"[autotune] default {default}", flush=True)
        res = r.autotune(hs.camera(), probe_spp=8, cache=None)
        print(f"[autotune] chose {res}", flush=True)
        assert res["spp_per_launch"] >= default["spp_per_launch"] // 2
        assert res["candidates"] >= 2 + 15
        probe = 4 * default["spp_per_launch"]

        def timed():
            r.render(hs.camera(), 1, default["spp_per_launch"], want_image=False)   # warm
            best = 1e30
            for _ in range(3):
                r.render(hs.camera(), 1, probe, want_image=False)
                best = min(best, r.stats()["render_ms"])
            return best

        t_win = timed()
        r.set_tuning(0, 0, 0)
        assert r.tuning() == default
        t_def = timed()
        print(f"[autotune] probe {probe} spp: winner {t_win:.1f} ms, default {t_def:.1f} ms", flush=True)
        assert t_win <= 1.05 * t_def


def test_autotune_with_guiding_keeps_capacity_and_network():
    """With guiding on, autotune searches workgroup size and LDS depth at the
    fixed queue capacity and does not train the network while probing."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(64, 48, spp=8)
    hs = HipScene(desc)
    with HipRenderer(64, 48, 8, 3, spp_per_launch=4) as r:
        r.upload_scene(hs)
        r.guide_enable(seed=3)
        r.render(hs.camera(), 1, 8, want_image=False)   # trains a few steps
        g0, w0 = r.guide_params()
        res = r.autotune(hs.camera(), probe_spp=4, cache=None)
        print(f"[autotune guided] {res}", flush=True)
        assert res["spp_per_launch"] == 4 and res["candidates"] == 1 + 15
        g1, w1 = r.guide_params()
        assert np.array_equal(g0, g1) and np.array_equal(w0, w1)
        r.render(hs.camera(), 1, 8, want_image=False)   # guided rendering and training still work
        assert r.stats()["paths"] == 64 * 48 * 8


def test_set_tuning_that_does_not_fit_keeps_previous_tuning():
    """A queue capacity the device cannot hold (2^30 paths at 292 B each is
    more than an MI355X's 288 GB) fails with XNH_E_ALLOC; the previous
    tuning is restored and renders as before."""
    from xenon_amd import HipRenderer, HipScene, native, scenes
    desc = scenes.cornell(1920, 1080, spp=2)
    hs = HipScene(desc)
    with HipRenderer(1920, 1080, 8, 3, spp_per_launch=1) as r:
        r.upload_scene(hs)
        r.render(hs.camera(), 1, 2, want_image=False)
        ref = r.read_accum()[0]
        before = r.tuning()
        ns = (1 << 30) // (1920 * 1080)
        with pytest.raises(native.XnhError) as e:
            r.set_tuning(128, 8, ns)
        print(f"[set_tuning] {e.value}", flush=True)
        assert e.value.code == -6 and "previous tuning is kept" in str(e.value)
        assert r.tuning() == before
        r.render(hs.camera(), 1, 2, want_image=False)
        assert np.array_equal(r.read_accum()[0].view(np.uint32), ref.view(np.uint32))
