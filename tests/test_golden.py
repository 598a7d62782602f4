"""The oracle and the scene generators reproduce the committed digests
(tests/golden/golden.json, written by tests/golden/make_golden.py)."""
import json
import os

import pytest

import make_golden as G

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "golden.json")) as f:
    GOLDEN = json.load(f)


def test_generator_digests():
    assert G.generators() == GOLDEN["generators"]


@pytest.fixture(scope="module")
def frames():
    return G.frames()


@pytest.mark.parametrize("name", sorted(GOLDEN["frames"]))
def test_oracle_frame_digests(frames, name):
    got, want = frames[name], GOLDEN["frames"][name]
    assert got["vis"] == want["vis"]
    assert got["depth"] == want["depth"]
    assert got["rgba"] == want["rgba"]
    if "stats" in want:
        assert got["stats"] == want["stats"]
