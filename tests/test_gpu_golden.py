"""Every layout's HIP traversal against the committed golden vectors: REF and PLAIN bit for bit,
CW8 identical visibility and closest hits within the tie rule."""
import glob
import os

import numpy as np
import pytest

import oracle
from helpers import classify_mismatch, compare_exact, gpu_trace
from renderlab_amd import rq

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "traverse_*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
@pytest.mark.parametrize("layout", ["ref", "plain", "cw8"])
def test_layout_reproduces_golden_vectors(path, layout):
    f = np.load(path)
    print(f"case: golden {os.path.basename(path)} {layout}", flush=True)
    scene = rq.Scene.from_gi_words(f["words"])
    scene.upload(("ref", "plain", "cw8"))
    rays = f["rays"].view(oracle.RAY_DTYPE)
    for mode in (0, 1, 2):
        h, v = gpu_trace(scene, layout, rays, mode)
        want = f[f"hits{mode}"].view(oracle.HIT_DTYPE)
        assert np.array_equal(v, f[f"vis{mode}"])
        if mode != 0:
            continue
        if layout == "cw8":
            genuine, _ = classify_mismatch(h, want)
            assert len(genuine) == 0
        else:
            assert len(compare_exact(h, want)) == 0
