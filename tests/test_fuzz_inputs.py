"""Inputs from outside the library -- reference-packed gi_geometry and
rl_geometry words, cooked cluster BVH ("RLB\\0") and near-field ("RNF\\0")
bytes -- mutated at random: every call either accepts them or refuses them
with InvalidInput, and words it accepts relayout (PLAIN, CW8, RS_CW8) and
decode their hits without error. Run against the sanitized host build
(`python -m renderlab_amd.build --sanitize`, RQ_LIBRARY) this also checks that
no parser reads or writes outside its input."""
import numpy as np
import pytest

from renderlab_amd import rq
from test_cooked import grid, reference_field, rnf_bytes

SPECIAL = np.array([0, 1, 2, 3, 15, 16, 31, 32, 0x7FFFFFFF, 0x80000000, 0x80000001, 0xFFFFFFFE, 0xFFFFFFFF,
                    0x3F800000, 0x7F800000, 0x7FC00000, 0xFF800000], np.uint32)


def mutate_words(g, base, k):
    w = base.copy()
    for _ in range(k):
        i = int(g.integers(len(w)))
        r = g.random()
        if r < 0.4:
            w[i] = SPECIAL[int(g.integers(len(SPECIAL)))]
        elif r < 0.7:
            w[i] = np.uint32(int(w[i]) + int(g.integers(-3, 4)) & 0xFFFFFFFF)
        elif r < 0.85:
            w[i] = np.uint32(g.integers(0, len(w) + 64))
        else:
            w[i] = np.uint32(g.integers(0, 1 << 32))
    if g.random() < 0.1:
        w = w[:int(g.integers(1, len(w)))]
    return w


def try_words(words, reflection):
    try:
        sc = rq.Scene.from_rl_words(words) if reflection else rq.Scene.from_gi_words(words)
    except rq.InvalidInput:
        return False
    # accepted words: the relayouts and the host hit decoder read them too
    for layout in (("rs_cw8",) if reflection else ("plain", "cw8")):
        try:
            sc.layout_image(layout, 0)
        except rq.InvalidInput:
            pass
    if not reflection:
        h = np.zeros(1, rq.HIT_DTYPE)
        h["flags"], h["inst"] = 1, 0
        sc.hit_record(h[0])
    return True


def small_scene(reflection):
    g = np.random.default_rng(1)
    meshes = []
    for k in range(3):
        p = g.normal(size=(30, 3)).astype(np.float32)
        meshes.append((p, np.arange(30, dtype=np.uint32), k % 2))
    inst = []
    for k in range(45):  # >= 40 instances: a top level
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = g.uniform(-5, 5, 3)
        inst.append(dict(model=m, mesh=k % 3, identity=k + 1, placement=k))
    return rq.Scene.build(meshes, inst, reflection=reflection).words


@pytest.mark.parametrize("reflection", [False, True])
def test_mutated_words_are_accepted_or_refused(reflection):
    base = small_scene(reflection)
    assert try_words(base, reflection)
    g = np.random.default_rng(2 + reflection)
    accepted = 0
    for trial in range(1500):
        accepted += try_words(mutate_words(g, base, int(g.integers(1, 6))), reflection)
    assert 0 < accepted < 1500  # some mutations are harmless, most are caught


def mutate_bytes(g, base, k):
    b = bytearray(base)
    for _ in range(k):
        i = int(g.integers(len(b)))
        r = g.random()
        if r < 0.5:
            b[i] = int(g.integers(256))
        elif r < 0.8:
            b[i] ^= 1 << int(g.integers(8))
        else:  # a whole little-endian word
            i -= i % 4
            if i + 4 <= len(b):
                b[i:i + 4] = int(SPECIAL[int(g.integers(len(SPECIAL)))]).to_bytes(4, "little")
    if g.random() < 0.1:
        b = b[:int(g.integers(1, len(b)))]
    elif g.random() < 0.05:
        b += bytes(int(g.integers(1, 9)))
    return bytes(b)


def test_mutated_cooked_files_are_accepted_or_refused():
    p, i = grid(6)
    rlb = rq.ClusterBvh.from_mesh(p, i, sah=True).encode()
    _, _, bvh, masks, transform = reference_field()
    rnf = rnf_bytes((1.0, 0.5, 0.01, 16), [7] * 32, [bvh.encode()], [(0, transform, masks)])
    g = np.random.default_rng(4)
    for decode, base in ((rq.ClusterBvh.decode, rlb), (rq.NearField.decode, rnf)):
        accepted = 0
        for trial in range(1500):
            data = mutate_bytes(g, base, int(g.integers(1, 5)))
            try:
                obj = decode(data)
            except rq.InvalidInput:
                continue
            accepted += 1
            obj.encode()  # an accepted object encodes again
        assert accepted < 1500
