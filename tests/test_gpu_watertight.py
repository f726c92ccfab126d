"""RQ_WATERTIGHT on the GPU: the watertight leaf test and conservative culling
equal the oracle's watertight restatement (pinned in test_watertight.py) --
every hit record bit for bit in the REF and PLAIN layouts, the same triangle
except on ties in CW8, identical visibility bits -- on closed meshes traced
from inside through every edge and vertex (where none may miss), on the exact
edge-sign cases, and on the box scene's 960x640 frame. Moller-Trumbore stays
the default: the same launch without the flag still equals the reference
oracle."""
import numpy as np
import pytest

import oracle
from helpers import classify_mismatch, compare_exact, gpu_trace, gpu_workload, vis_bits
from renderlab_amd import rq
from test_watertight import from_centre, icosphere

pytestmark = pytest.mark.gpu
WT = rq.WATERTIGHT
LAYOUTS = ("ref", "plain", "cw8")


def check(scene, rays, layouts=LAYOUTS):
    want, _, _ = oracle.traverse(scene.words, rays, WT)
    _, want_vis, _ = oracle.traverse(scene.words, rays, WT | rq.ANY, want_hits=False, want_vis=True)
    for layout in layouts:
        got, _ = gpu_trace(scene, layout, rays, WT)
        if layout == "cw8":
            bad, _ = classify_mismatch(got, want)
            assert len(bad) == 0, layout
        else:
            assert len(compare_exact(got, want)) == 0, layout
        _, vis = gpu_trace(scene, layout, rays, WT | rq.ANY)
        assert np.array_equal(vis_bits(vis, len(rays)), vis_bits(want_vis, len(rays))), layout
    return want


@pytest.mark.parametrize("centre", [(0.0, 0.0, 0.0), (1000.3, -700.7, 512.1)])
def test_closed_meshes_from_inside_never_miss(centre):
    for seed in range(3):
        pos, idx = icosphere(3, seed, 3.0, np.array(centre))
        scene = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4, dtype=np.float32), mesh=0)])
        scene.upload(LAYOUTS)
        rays = from_centre(pos, idx, centre)
        print(f"watertight: icosphere seed {seed} at {centre}, {len(rays)} rays", flush=True)
        want = check(scene, rays)
        assert ((want["flags"] & 1) == 1).all()
        for layout in LAYOUTS:
            got, _ = gpu_trace(scene, layout, rays, WT)
            assert ((got["flags"] & 1) == 1).all(), layout


def test_box_scene_frame_in_watertight_mode():
    scene = rq.Scene.generate(0, 1)
    scene.upload(LAYOUTS)
    _, prim, hits, shadow, _ = gpu_workload(scene, 0, 1, 960, 640, 0, "ref")
    print(f"watertight: box scene, {len(prim)} primary and {len(shadow)} shadow rays", flush=True)
    check(scene, prim)
    check(scene, shadow)
    # without the flag the same launch is the reference traversal
    want, _, _ = oracle.traverse(scene.words, prim, 0)
    got, _ = gpu_trace(scene, "plain", prim, rq.CLOSEST)
    assert len(compare_exact(got, want)) == 0


def test_instanced_scene_with_a_top_level_in_watertight_mode():
    g = np.random.default_rng(2)
    meshes = [icosphere(2, s, 1.0, np.zeros(3)) for s in range(3)]
    inst = []
    for k in range(60):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = g.uniform(-20, 20, 3)
        m[:3, :3] *= g.uniform(0.5, 2.0)
        inst.append(dict(model=m, mesh=k % 3, flags=int(k % 7 == 0)))
    scene = rq.Scene.build([(p, i, 1) for p, i in meshes], inst)
    assert scene.words[1] != 0  # the top level is on
    scene.upload(LAYOUTS)
    rays = np.zeros(20000, rq.RAY_DTYPE)
    rays["origin"] = g.uniform(-30, 30, (20000, 3))
    rays["direction"] = g.normal(size=(20000, 3))
    rays["tmax"] = 1e20
    check(scene, rays)
