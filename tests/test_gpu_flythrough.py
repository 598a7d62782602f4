"""GPU: the combined workload at reduced size -- terrain plus instanced
icospheres, every frame one batched launch of a main view and six cube-map
faces (views of one scene), learned single-pass binning with exact
fallback.  Every frame must equal the exact binner's frame bit for bit
(vis, depth and RGBA8), and the main view of frame 0 the oracle's."""
import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S

pytestmark = pytest.mark.gpu


def make_views(fly, tile=16):
    owner = R.Renderer(*fly.main, tile=tile)
    owner.add_scene(fly.scene)
    return [owner] + [owner.view(fly.face, fly.face, tile=tile) for _ in range(6)]


def close_views(vs):
    for v in vs[1:]:
        v.close()
    vs[0].close()


def test_flythrough_learned_batches_equal_the_exact_binner():
    fly = S.config5(frames=10, n=160, spheres=8, subdiv=3, tex_size=256, main=(480, 272), face=96)
    train = S.Flythrough(fly.scene, S.camera_path(S.SEED_BASE + 0x505, 6), fly.main, fly.face)
    learned, exact = make_views(fly), make_views(fly)
    models = [R.BinModel(v, hidden=64, seed=20 + k) for k, v in enumerate(learned)]
    try:
        for m, v in zip(models, learned):
            m.attach(v, 1.1, 2)
        for f in range(len(train.path)):
            R.render_batch([(v, cam, fly.opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
                            for v, (cam, _, _) in zip(learned, train.views(f))])
            if f > 0:
                for m, v in zip(models, learned):
                    m.record(v)
        for m in models:
            m.train(epochs=4, batch=1024, lr=2e-3, tau=0.95)
        overflow = 0
        for f in range(len(fly.path)):
            views = fly.views(f)
            R.render_batch([(v, cam, fly.opts, R.BIN_LEARNED, 0) for v, (cam, _, _) in zip(learned, views)])
            R.render_batch([(v, cam, fly.opts, R.BIN_EXACT, 0) for v, (cam, _, _) in zip(exact, views)])
            for k, (a, b) in enumerate(zip(learned, exact)):
                ga, gb = a.readback(), b.readback()
                for name, x, y in zip(("vis", "depth", "rgba"), ga, gb):
                    bad = np.count_nonzero(x.view(np.uint32) != y.view(np.uint32))
                    assert bad == 0, f"frame {f} view {k}: {name} differs at {bad} pixels"
                overflow += a.stats()["overflow_tiles"]
            if f == 0:
                cam, W, H = views[0]
                ovis, odepth, _, _ = O.render(fly.scene, cam, fly.opts, W, H, threads=8, want_rgba=False)
                vis, depth, _ = learned[0].readback()
                np.testing.assert_array_equal(vis, ovis)
                np.testing.assert_array_equal(depth.view(np.uint32), odepth.view(np.uint32))
        print("overflow tiles over the path:", overflow)
    finally:
        for m in models:
            m.close()
        close_views(learned)
        close_views(exact)
