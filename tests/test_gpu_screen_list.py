"""GPU: triangles on the screen-wide list (spanning more than big_threshold
tiles) against the oracle, vis and depth bit for bit.  The flythrough
scene's cube-map faces look along the terrain from a camera standing on it,
so many triangles cross the plane w = 0: the reference has no near clipping
and projects them to slivers with screen coordinates beyond 1e4 and depths
far outside [0, 1], of which a few pixels survive the depth range test and
win.  Every tile size is run with a screen-list threshold of 8 tiles
and forced to hold nearly every triangle (threshold 1, no setup-time
raster), so that the per-tile front-to-back pass and, where more triangles
survive in a tile than it holds, the pass over the whole list in stored
order both run."""
import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fly():
    return S.config5(frames=4, n=160, spheres=8, subdiv=3, tex_size=256, main=(480, 272), face=128)


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("big,micro", [(8, 16), (1, 0)])
def test_cube_faces_match_the_oracle(fly, tile, big, micro):
    owner = R.Renderer(*fly.main, tile=tile, big_threshold=big)
    owner.set_tuning(micro, 16)
    owner.add_scene(fly.scene)
    face = owner.view(fly.face, fly.face, tile=tile, big_threshold=big)
    face.set_tuning(micro, 16)
    try:
        big_seen = 0
        for f in (0, 2, 3):
            for k, (cam, W, H) in enumerate(fly.views(f)):
                r = owner if k == 0 else face
                r.render(cam, fly.opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
                vis, depth, _ = r.readback(rgba=False)
                big_seen += r.stats()["num_big"]
                ovis, odepth, _, _ = O.render(fly.scene, cam, fly.opts, W, H, threads=8, want_rgba=False)
                bad = np.count_nonzero(vis != ovis)
                assert bad == 0, f"frame {f} view {k}: vis differs at {bad} pixels"
                bad = np.count_nonzero(depth.view(np.uint32) != odepth.view(np.uint32))
                assert bad == 0, f"frame {f} view {k}: depth differs at {bad} pixels"
        assert big_seen > 0
    finally:
        face.close()
        owner.close()
