"""GPU parity: every layout's HIP traversal against the oracle on the same rays.

REF and PLAIN must reproduce the oracle's 48-byte hit records bit for bit
(same boxes, same visit order, same arithmetic). CW8 visits children in a
different order over different (outward-quantized) boxes, so its hit
triangle may differ only on an equal-distance tie (|dt|/t <= 1e-5), with t
and barycentrics within 1e-5 relative and identical visibility bits.
"""
import numpy as np
import pytest

import oracle
from helpers import chain_words, classify_mismatch, compare_exact, gpu_trace, gpu_workload, vis_bits
from renderlab_amd import rq

pytestmark = pytest.mark.gpu

W, H = 240, 160


@pytest.fixture(scope="module")
def scene_c0():
    s = rq.Scene.generate(0, 1)
    s.upload(("ref", "plain", "cw8"))
    return s


@pytest.fixture(scope="module")
def workload_c0(scene_c0):
    print(f"workload: C0 seed 1 {W}x{H} P+S+GI", flush=True)
    return gpu_workload(scene_c0, 0, 1, W, H, 1, layout="ref")


def test_primary_rays_match_oracle_bits(workload_c0):
    inv, prim, _, _, _ = workload_c0
    want = oracle.primary_rays(inv, W, H)
    assert (prim.view(np.uint32) == want.view(np.uint32)).all()


def test_tiled_primary_rays_are_the_pixels_in_tile_order(workload_c0):
    """rq_gen_primary_tiled: ray i is pixel i % 64 of 8x8 tile i / 64 (tiles and pixels
    row-major), bit for bit the oracle's ray of that pixel; sizes not multiples of 8 refused."""
    inv, _, _, _, _ = workload_c0
    want = oracle.primary_rays(inv, W, H)
    d = rq.DeviceBuffer(W * H * 32)
    rq.gen_primary(inv, W, H, d, tiled=True)
    rq.synchronize()
    got = d.download(rq.RAY_DTYPE, W * H)
    i = np.arange(W * H)
    t, k = i // 64, i % 64
    px = (t % (W // 8)) * 8 + k % 8
    py = (t // (W // 8)) * 8 + k // 8
    assert (got.view(np.uint32).reshape(-1, 8) == want[py * W + px].view(np.uint32).reshape(-1, 8)).all()
    with pytest.raises(rq.InvalidInput):
        rq.gen_primary(inv, W - 4, H, d, tiled=True)


@pytest.mark.parametrize("layout", ["ref", "plain"])
def test_closest_bit_exact(scene_c0, workload_c0, layout):
    _, prim, _, _, gis = workload_c0
    words = scene_c0.words
    for name, rays in (("primary", prim), ("gi", gis[0])):
        print(f"closest {layout} {name} n={len(rays)}", flush=True)
        g, _ = gpu_trace(scene_c0, layout, rays, rq.CLOSEST)
        o, _, _ = oracle.traverse(words, rays, 0)
        bad = compare_exact(g, o)
        assert bad.size == 0, f"{layout}/{name}: {bad.size} records differ, first {bad[:5]}"


@pytest.mark.parametrize("layout", ["ref", "plain", "cw8"])
@pytest.mark.parametrize("mode", [1, 2])
def test_visibility_bits(scene_c0, workload_c0, layout, mode):
    _, _, _, shadow, _ = workload_c0
    words = scene_c0.words
    print(f"any-hit {layout} mode {mode} n={len(shadow)}", flush=True)
    _, gv = gpu_trace(scene_c0, layout, shadow, mode)
    _, ov, _ = oracle.traverse(words, shadow, mode, want_hits=False, want_vis=True)
    assert (gv == ov).all()


def test_cw8_closest_within_tolerance(scene_c0, workload_c0):
    _, prim, _, _, gis = workload_c0
    words = scene_c0.words
    for name, rays in (("primary", prim), ("gi", gis[0])):
        print(f"closest cw8 {name} n={len(rays)}", flush=True)
        g, _ = gpu_trace(scene_c0, "cw8", rays, rq.CLOSEST)
        o, _, _ = oracle.traverse(words, rays, 0)
        bad, ties = classify_mismatch(g, o)
        assert bad.size == 0, f"cw8/{name}: {bad.size} non-tie mismatches, first {bad[:5]}"
        both = ((g["flags"] & 1) == 1) & ((o["flags"] & 1) == 1)
        rel = np.abs(g["t"][both].astype(np.float64) - o["t"][both]) / np.abs(o["t"][both])
        assert rel.max(initial=0.0) <= 1e-5
        same = both & (g["inst"] == o["inst"]) & (g["slot"] == o["slot"])
        assert (g["u"][same] == o["u"][same]).all() and (g["v"][same] == o["v"][same]).all()


def test_cw8_speculation_changes_order_only(scene_c0, workload_c0):
    """Node-step speculation (lanes with triangle pairs pending join the wave's node steps) only
    reorders each lane's visits: visibility bits are identical with it off and on, and closest
    hits differ only between equal-distance candidates; a hit on the same triangle is the same
    record."""
    _, prim, _, shadow, gis = workload_c0
    base = [rq.get_tuning("cw8", m) for m in (0, 1)]
    out = {}
    try:
        for spec in (0, 1):
            for m, t in enumerate(base):
                rq.set_tuning("cw8", m, t.block, t.stack, t.refill, t.grid_per_cu, speculate=spec, drain=t.drain)
            print(f"cw8 speculate={spec}", flush=True)
            out[spec] = ([gpu_trace(scene_c0, "cw8", r, rq.CLOSEST)[0] for r in (prim, gis[0])],
                         gpu_trace(scene_c0, "cw8", shadow, 1)[1])
    finally:
        for m, t in enumerate(base):
            rq.set_tuning("cw8", m, t.block, t.stack, t.refill, t.grid_per_cu, speculate=t.speculate, drain=t.drain)
    assert (out[0][1] == out[1][1]).all()
    for g0, g1 in zip(out[0][0], out[1][0]):
        bad, _ = classify_mismatch(g1, g0)
        assert bad.size == 0, f"{bad.size} non-tie differences, first {bad[:5]}"
        same = (g0["inst"] == g1["inst"]) & (g0["slot"] == g1["slot"])
        for k in ("t", "u", "v", "flags"):
            assert (g0[k][same] == g1[k][same]).all(), k


def test_secondary_rays_match_oracle(workload_c0):
    inv, prim, hits, shadow, gis = workload_c0
    masks, rays = oracle.secondary_rays(prim, hits, np.array([0.3, 0.8, 0.2], np.float32) /
                                        np.float32(np.linalg.norm([0.3, 0.8, 0.2])), 1, 1)
    want_s = rays[(masks & 1) == 1, 0]
    want_g = rays[(masks & 2) == 2, 1]
    assert len(shadow) == len(want_s) and len(gis[0]) == len(want_g)
    assert (shadow.view(np.uint32) == want_s.view(np.uint32)).all()
    # GI directions go through cosf/sinf: device and host libm may differ in the last bit.
    assert np.allclose(gis[0]["direction"], want_g["direction"], rtol=0, atol=2e-6)
    assert (gis[0]["origin"] == want_g["origin"]).all()


def test_fp64_verification(scene_c0, workload_c0):
    _, prim, _, _, _ = workload_c0
    dr = rq.DeviceBuffer.from_array(prim)
    dh = rq.DeviceBuffer(len(prim) * 48)
    scene_c0.trace("plain", dr, len(prim), rq.CLOSEST, hits=dh)
    v = scene_c0.verify_fp64(dr, dh, len(prim))
    host = dh.download(rq.HIT_DTYPE, len(prim))
    ref = oracle.reintersect_f64(scene_c0.words, prim, host)
    hit = (host["flags"] & 1) == 1
    rel = np.abs(host["t"][hit] - ref["t"][hit]) / np.abs(ref["t"][hit])
    assert v["checked"] == hit.sum()
    assert abs(v["max_t_rel"] - rel.max()) <= 1e-12 + 1e-9 * rel.max()
    assert v["max_t_rel"] < 1e-5
    assert v["over_1e5"] == int((rel > 1e-5).sum())
    assert 0 < v["conditioned"] <= v["checked"]
    assert v["max_ulps_conditioned"] < 16


def test_fp64_verification_gi_conditioned(scene_c0, workload_c0):
    """Incoherent GI hits: away from grazing and slivers, the f32 hit distance
    is within a few f32 roundings of the numbers it is computed from
    (the f32 world-to-object transform of a ray far from a small instance
    cancels digits, so a relative bound on t alone does not hold)."""
    _, _, _, _, gis = workload_c0
    g = gis[0]
    dr = rq.DeviceBuffer.from_array(g)
    dh = rq.DeviceBuffer(len(g) * 48)
    scene_c0.trace("cw8", dr, len(g), rq.CLOSEST, hits=dh)
    v = scene_c0.verify_fp64(dr, dh, len(g))
    print(f"fp64 gi: {v}", flush=True)
    assert v["conditioned"] > 0.5 * v["checked"]
    assert v["max_ulps_conditioned"] < 16


def test_cw8_deep_chain_against_the_flat_answer():
    """A mesh whose 4-wide tree is a 100-node chain (one that does not hold triangle 0, so the
    relayout keeps the chain's shape) collapses into a 15-level CW8 tree: the CW8 stack (7 entries
    a level, no bound check) holds it, and every ray straight down onto a triangle of the tree finds
    that triangle at t = 1, as the tree-free flat loop does (the reference's bounded stack drops
    the chain's deeper nodes); the ray onto triangle 0 finds nothing."""
    scene = rq.Scene.from_gi_words(chain_words(100, hole=True))
    scene.upload(("cw8",))
    rays = np.zeros(101, rq.RAY_DTYPE)
    rays["origin"] = np.stack([np.arange(101) + 0.1, np.full(101, 0.1), np.ones(101)], axis=1)
    rays["direction"] = [0.0, 0.0, -1.0]
    rays["tmax"] = 10.0
    print("cw8 chain(100) n=101", flush=True)
    g, gv = gpu_trace(scene, "cw8", rays, rq.CLOSEST)
    f = oracle.flat_f32(scene.words, rays, 0)
    assert (g["slot"][1:] == f["slot"][1:]).all() and (g["t"][1:] == f["t"][1:]).all()
    hit = (g["flags"] & 1) == 1
    assert not hit[0] and hit[1:].all()  # triangle 0 is in no leaf
    assert (g["slot"][1:] == np.arange(1, 101)).all() and (g["t"][1:] == 1.0).all()
    _, gv = gpu_trace(scene, "cw8", rays, 1)
    assert (vis_bits(gv, 101) == ~hit).all()  # blocked but for ray 0


@pytest.mark.parametrize("count", [0, 1, 63, 64, 65, 1000])
def test_small_launches_match_oracle(scene_c0, workload_c0, count):
    """Launches smaller than a wave or than the grid's first handout (each wave starts on its own
    64 rays without the cursor): REF and PLAIN records bit for bit, CW8 within its tie rule."""
    _, _, _, _, gis = workload_c0
    rays = gis[0][:count]
    o, ov, _ = oracle.traverse(scene_c0.words, rays, 0, want_vis=False) if count else (None, None, None)
    for layout in ("ref", "plain", "cw8"):
        print(f"closest {layout} n={count}", flush=True)
        g, _ = gpu_trace(scene_c0, layout, rays, rq.CLOSEST)
        if not count:
            continue
        if layout == "cw8":
            bad, _ = classify_mismatch(g, o)
        else:
            bad = compare_exact(g, o)
        assert bad.size == 0, f"{layout}: {bad.size} records differ, first {bad[:5]}"



def test_cw8_without_a_top_level_in_the_words():
    """Words with no top level (forced off with 50 instances; a scene of 2 instances): CW8 builds one
    all the same, since it tests no record box at instance entry. Closest hits agree with the
    oracle within the tie rule, PLAIN's records equal it bit for bit, visibility bits are equal."""
    from test_reflection import random_reflection_scene, random_rays
    meshes, inst, g = random_reflection_scene(6)
    rays = random_rays(g, 4000)
    for name, sc in (("50 instances, flat words", rq.Scene.build(meshes, inst, top_level=0)),
                     ("2 instances", rq.Scene.build(meshes, inst[:2]))):
        assert sc.words[1] == 0
        sc.upload(("plain", "cw8"))
        assert sc.layout_image("cw8", 5)[0] == 1
        o, _, _ = oracle.traverse(sc.words, rays, 0)
        assert ((o["flags"] & 1) == 1).sum() > 50
        print(f"case: cw8 and plain, {name} n={len(rays)}", flush=True)
        p, _ = gpu_trace(sc, "plain", rays, rq.CLOSEST)
        assert compare_exact(p, o).size == 0
        c, _ = gpu_trace(sc, "cw8", rays, rq.CLOSEST)
        bad, _ = classify_mismatch(c, o)
        assert bad.size == 0, f"{name}: {bad.size} non-tie mismatches, first {bad[:5]}"
        _, ov1, _ = oracle.traverse(sc.words, rays, 1, want_hits=False, want_vis=True)
        _, cv = gpu_trace(sc, "cw8", rays, 1)
        assert np.array_equal(cv, ov1)


def test_axis_parallel_rays_with_signed_zeros(scene_c0):
    """Rays along the axes whose other direction components are +0 or -0 (the 8-wide box test takes
    the reciprocal of each component, clamped to 1e30 with the component's sign, and picks near and
    far planes by that sign): CW8 agrees with the oracle within the tie rule, PLAIN bit for bit, and
    visibility bits are equal."""
    g = np.random.default_rng(11)
    n = 6000
    rays = np.zeros(n, rq.RAY_DTYPE)
    rays["origin"] = g.uniform(-60, 60, (n, 3))
    axis = g.integers(0, 3, n)
    d = np.zeros((n, 3), np.float32)
    d[np.arange(n), axis] = np.where(g.random(n) < 0.5, -1.0, 1.0)
    zero_sign = g.random((n, 3)) < 0.5
    d = np.where((d == 0) & zero_sign, np.float32(-0.0), d)
    rays["direction"] = d
    rays["tmax"] = 1e20
    assert (np.signbit(rays["direction"]) & (rays["direction"] == 0)).any()
    o, _, _ = oracle.traverse(scene_c0.words, rays, 0)
    assert ((o["flags"] & 1) == 1).sum() > 200
    print(f"case: axis-parallel rays n={n}", flush=True)
    p, _ = gpu_trace(scene_c0, "plain", rays, rq.CLOSEST)
    assert compare_exact(p, o).size == 0
    c, _ = gpu_trace(scene_c0, "cw8", rays, rq.CLOSEST)
    bad, _ = classify_mismatch(c, o)
    assert bad.size == 0, f"{bad.size} non-tie mismatches, first {bad[:5]}"
    _, ov, _ = oracle.traverse(scene_c0.words, rays, 1, want_hits=False, want_vis=True)
    _, cv = gpu_trace(scene_c0, "cw8", rays, 1)
    assert np.array_equal(cv, ov)
