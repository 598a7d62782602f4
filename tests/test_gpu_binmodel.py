"""The learned bin-capacity model on the GPU: forward and loss gradient
against the CPU oracle, training on exact counts, and what it gains over the
previous-frame heuristic on a disjoint camera path (output unchanged)."""
import numpy as np
import pytest

import oracle as O
import rast as R
import scenes as S

pytestmark = pytest.mark.gpu

F = 8


@pytest.fixture(scope="module")
def renderer():
    r = R.Renderer(64, 64)
    yield r
    r.close()


@pytest.mark.parametrize("hidden", [16, 32, 64])
def test_forward_matches_oracle(renderer, hidden):
    m = R.BinModel(renderer, hidden=hidden, seed=5)
    rng = np.random.default_rng(hidden)
    p = rng.normal(0, 0.3, m.num_params).astype(np.float32)
    m.set_params(p)
    x = rng.normal(0, 1, (1000, F)).astype(np.float32)
    got = m.forward(x)
    want = O.mlp_forward(p, F, hidden, x)
    m.close()
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=2e-4)


@pytest.mark.parametrize("hidden", [16, 64])
@pytest.mark.parametrize("tau", [0.5, 0.9])
def test_loss_gradient_matches_oracle(renderer, hidden, tau):
    m = R.BinModel(renderer, hidden=hidden, seed=6)
    rng = np.random.default_rng(7)
    p = rng.normal(0, 0.3, m.num_params).astype(np.float32)
    m.set_params(p)
    n = 3000
    x = rng.normal(0, 1, (n, F)).astype(np.float32)
    y = rng.normal(2, 1, n).astype(np.float32)
    loss, g = m.loss_grad(x, y, tau)
    oloss, og = O.mlp_loss_grad(p, F, hidden, x, y, tau)
    m.close()
    assert abs(loss - oloss) < 1e-4 * max(1.0, abs(oloss))
    scale = np.abs(og).max()
    np.testing.assert_allclose(g, og, rtol=2e-3, atol=2e-4 * scale)


def small_terrain():
    return S.config3(n=160, tex_size=128, width=640, height=352)


def record_and_train(r, w, frames, hidden=64, epochs=40):
    m = R.BinModel(r, hidden=hidden, seed=11)
    m.attach(r, 1.25, 4)
    for i, cam in enumerate(frames):
        r.render(cam, w.opts, R.BIN_EXACT, R.FLAG_NO_SHADE)
        if i > 0:
            m.record(r)
    return m


def run_path(r, w, path, policy):
    stats = []
    for cam in path:
        r.render(cam, w.opts, policy, R.FLAG_NO_SHADE)
        stats.append(r.stats())
    return stats


def test_training_lowers_the_loss_and_learned_binning_beats_previous_frame_counts():
    w = small_terrain()
    train, test = S.train_test_paths(60, 30)
    with R.Renderer(w.width, w.height) as r:
        r.set_tuning(0, 16)  # bin every triangle: the queues carry the whole frame
        r.add_scene(w.scene)
        m = record_and_train(r, w, train)
        loss0 = m.train(epochs=1, batch=4096, lr=2e-3, tau=0.9)
        loss1 = m.train(epochs=60, batch=4096, lr=2e-3, tau=0.9)
        assert loss1 < 0.7 * loss0, (loss0, loss1)
        # exact reference numbers on the test path
        exact = run_path(r, w, test, R.BIN_EXACT)
        m.detach(r)
        heur = run_path(r, w, test, R.BIN_PREVFRAME)
        m.attach(r, 1.25, 4)
        learned = run_path(r, w, test, R.BIN_LEARNED)
        m.close()
    pairs = sum(s["num_bin_pairs"] for s in exact[1:])
    h_ovf = sum(s["overflow_entries"] for s in heur[1:])
    l_ovf = sum(s["overflow_entries"] for s in learned[1:])
    h_cap = sum(s["capacity_total"] for s in heur[1:])
    l_cap = sum(s["capacity_total"] for s in learned[1:])
    print(f"pairs {pairs} heuristic overflow {h_ovf} cap {h_cap} | learned overflow {l_ovf} cap {l_cap}")
    assert l_ovf < h_ovf            # fewer entries fall back to exact binning
    assert l_cap < 3 * pairs         # without buying it with a huge allocation


def test_learned_frames_are_bit_identical_to_exact(renderer):
    w = small_terrain()
    train, test = S.train_test_paths(12, 8)
    with R.Renderer(w.width, w.height) as ex, R.Renderer(w.width, w.height) as lr:
        ex.add_scene(w.scene)
        lr.add_scene(w.scene)
        m = record_and_train(lr, w, train)
        m.train(epochs=5, batch=2048)
        for cam in test:
            ex.render(cam, w.opts, R.BIN_EXACT)
            lr.render(cam, w.opts, R.BIN_LEARNED)
            a, b = ex.readback(), lr.readback()
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
        m.close()


@pytest.mark.parametrize("hidden", [16, 32, 64])
def test_frame_capacities_match_the_oracle_forward(hidden):
    """The frame-time capacity kernel (matrix-core layers, 16 tiles per wave)
    against the oracle's forward pass.  The first layer's weights are zero
    except on the two tile-position features, so each tile's prediction is a
    function of its position alone; capacity = ceil(expm1(out) * margin) +
    slack (at most the triangle count), exact except where the value sits
    within 1e-4 of an integer."""
    w = small_terrain()
    tile, margin, slack = 16, 1.25, 4
    rng = np.random.default_rng(100 + hidden)
    with R.Renderer(w.width, w.height, tile=tile) as r:
        r.add_scene(w.scene)
        m = R.BinModel(r, hidden=hidden, seed=9)
        p = rng.normal(0, 0.4, m.num_params).astype(np.float32)
        W1 = p[:hidden * F].reshape(hidden, F)
        W1[:, :6] = 0.0
        W1[:, 6:] *= 6.0  # position only, steep enough to vary over the screen
        o3 = hidden * F + hidden + hidden * hidden + hidden
        p[o3:o3 + hidden] *= 2.0
        p[-1] = 2.5
        m.set_params(p)
        m.attach(r, margin, slack)
        cam = S.train_test_paths(2, 1)[1][0]
        r.render(cam, w.opts, R.BIN_EXACT)
        # the first single-pass frame may find its capacities cut to the queue
        # storage; reading it grows the storage, and the second has room
        r.render(cam, w.opts, R.BIN_LEARNED)
        r.sync()
        r.render(cam, w.opts, R.BIN_LEARNED)
        caps = r.debug_read(R.DEBUG_TILE_CAPS, np.uint32)
        m.close()
    tx_n, ty_n = -(-w.width // tile), -(-w.height // tile)
    assert caps.size >= tx_n * ty_n
    caps = caps[:tx_n * ty_n]
    t = np.arange(tx_n * ty_n)
    x = np.zeros((t.size, F), np.float32)
    x[:, 6] = ((t // tx_n).astype(np.float32) + np.float32(0.5)) / np.float32(ty_n)
    x[:, 7] = ((t % tx_n).astype(np.float32) + np.float32(0.5)) / np.float32(tx_n)
    out = O.mlp_forward(p, F, hidden, x).astype(np.float64)
    v = np.expm1(np.minimum(out, 20.0)) * margin
    want = np.minimum(np.maximum(np.ceil(v), 0) + slack, w.scene.num_triangles)
    near = np.abs(v - np.round(v)) < 1e-4 * np.maximum(1.0, np.abs(v))
    diff = caps.astype(np.int64) - want.astype(np.int64)
    assert np.all(diff[~near] == 0), f"{np.count_nonzero(diff[~near])} tiles differ"
    assert np.all(np.abs(diff[near]) <= 1)
    assert len(np.unique(caps)) > 20  # the predictions vary over the screen


@pytest.mark.parametrize("policy", [R.BIN_LEARNED, R.BIN_PREVFRAME])
def test_saturated_capacities_still_give_the_exact_frames(policy):
    """A model whose output saturates (expm1(20) x margin 64 entries per tile,
    far more than the triangle count, summed over the tiles far past 2^32)
    and the largest slack: every capacity is clamped to the triangle count,
    their sum to the queue storage (the tiles past it keep no room and fall
    back to exact binning), and the frames equal the exact binner's."""
    w = small_terrain()
    _, path = S.train_test_paths(2, 3)
    with R.Renderer(w.width, w.height, tile=8) as ex, R.Renderer(w.width, w.height, tile=8) as sp:
        ex.add_scene(w.scene)
        sp.add_scene(w.scene)
        sp.set_tuning(0, 16)  # bin every triangle
        m = R.BinModel(sp, hidden=16, seed=4)
        p = np.zeros(m.num_params, np.float32)
        p[-1] = 30.0  # b3: every prediction saturates
        m.set_params(p)
        m.attach(sp, 64.0, 1 << 20)
        for cam in path:
            ex.render(cam, w.opts, R.BIN_EXACT)
            sp.render(cam, w.opts, policy)
            a, b = ex.readback(), sp.readback()
            for name, x, y in zip(("vis", "depth", "rgba"), a, b):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
        caps = sp.debug_read(R.DEBUG_TILE_CAPS, np.uint32)
        assert caps.max() <= w.scene.num_triangles
        m.close()
