"""The extraction walk on the device (svxg_extract_structure) against the oracle's sequential walk
(orc_extract_structure: World::Impl::extract's walk and append_nodes, world.cpp:425-620, 739-865):
members, their nodes, the nodes, identities, section heads, faces, bonds and bond identities, bit for
bit, and the walk's flags - over broken faces and damage, the node limit, the radius and owners,
clusters of both keys, the box's outside as the unknown world, a lone fragment, negative coordinates
and a large block."""
import numpy as np
import pytest

import oracle
from opus_destruct_2_amd import voxels
from opus_destruct_2_amd.abi import extract_options
from tests import walk_scenes as W

pytestmark = pytest.mark.gpu

KEYS = ("members", "member_node", "nodes", "node_ident", "heads", "faces", "bonds", "bond_ident")
FLAGS = ("cell", "truncated", "any_support", "touched_unloaded")


def _both(sc, fr, seed, mats, ctx, **kw):
    args = (sc["vox"], sc["frag"], sc["lo"], sc["h"], fr, seed, mats)
    o = oracle.extract_structure(*args, broken=sc["broken"], damage=sc["damage"], opts=extract_options(**kw))
    d = voxels.extract_structure(*args, broken=sc["broken"], damage=sc["damage"], opts=extract_options(**kw), ctx=ctx)
    return o, d


def _first_difference(a, b):
    """(index, field) of the first element where a and b differ in bytes, or None."""
    if len(a) == 0:
        return None
    ra = a.view(np.uint8).reshape(len(a), -1)
    rb = b.view(np.uint8).reshape(len(b), -1)
    rows = np.flatnonzero(np.any(ra != rb, 1))
    if not len(rows):
        return None
    i = int(rows[0])
    fields = [f for f in (a.dtype.names or ()) if a[f][i].tobytes() != b[f][i].tobytes()]
    return i, fields, len(rows)


def _same(o, d):
    for k in FLAGS:
        assert o[k] == d[k], k
    for k in KEYS:
        assert o[k].dtype == d[k].dtype and len(o[k]) == len(d[k]), (k, len(o[k]), len(d[k]))
        diff = _first_difference(o[k], d[k])
        if diff is not None:
            i, fields, count = diff
            lo = max(0, i - 2)
            raise AssertionError(f"{k}: {count} rows differ, the first at {i} (fields {fields}); oracle "
                                 f"{o[k][lo:i + 3]!r:.1500}; device {d[k][lo:i + 3]!r:.1500}")


def _seed(sc, z=3, y=10, x=10):
    return int(sc["frag"][z, y, x])


@pytest.mark.parametrize("broken_share,damage", [(0.0, False), (0.08, True)])
def test_walk_is_the_oracles(broken_share, damage, mats, ctx):
    sc = W.scene(broken_share=broken_share, damage=damage, seed=4)
    fr = W.frags_of(sc, mats)
    o, d = _both(sc, fr, _seed(sc), mats, ctx)
    assert len(o["members"]) > 100 and len(o["bonds"]) > 300
    _same(o, d)


def test_node_limit(mats, ctx):
    sc = W.scene(seed=6, broken_share=0.02)
    fr = W.frags_of(sc, mats)
    o, d = _both(sc, fr, _seed(sc), mats, ctx, max_nodes=300)
    assert o["truncated"] == 1 and len(o["members"]) == 300
    _same(o, d)


def test_radius_and_owners(mats, ctx):
    sc = W.scene(seed=7)
    fr = W.frags_of(sc, mats)
    seed = _seed(sc)
    dist = np.linalg.norm(fr["com"] - fr["com"][seed], axis=1)
    fr["owner"][np.flatnonzero((dist > 0.6) & (dist < 1.2))[:4]] = 9
    fr["owner"][np.flatnonzero((dist > 2.0) & (dist < 2.8))[:25]] = 4
    o, d = _both(sc, fr, seed, mats, ctx, max_radius=2.0, fresh_from=8)
    assert o["truncated"] == 1
    _same(o, d)


@pytest.mark.parametrize("limit,cubes", [(200, 0), (200, 1), (150, 0), (150, 1)])
def test_clusters(limit, cubes, mats, ctx):
    sc = W.scene(shape=(48, 40, 24), seed=8, broken_share=0.03)
    fr = W.frags_of(sc, mats)
    o, d = _both(sc, fr, _seed(sc), mats, ctx, cluster_nodes=limit, cluster_cubes=cubes)
    assert o["cell"] in (8, 16) and len(o["nodes"]) < len(o["members"])
    _same(o, d)


def test_outside_holds_and_a_free_piece(mats, ctx):
    sc = W.scene(shape=(36, 20, 12), seed=9, voids=0.0)
    sc["vox"] = sc["vox"][1:].copy()
    sc["frag"] = sc["frag"][1:].copy()
    sc["lo"] = sc["lo"] + np.array([0, 0, 1], np.int32)
    fr = W.frags_of(sc, mats)
    seed = _seed(sc, z=2)
    o, d = _both(sc, fr, seed, mats, ctx)
    assert o["any_support"] == 0
    _same(o, d)
    o, d = _both(sc, fr, seed, mats, ctx, outside=1)
    assert o["any_support"] == 1 and o["touched_unloaded"] == 1
    _same(o, d)


def test_lone_fragment(mats, ctx):
    sc = W.scene(shape=(20, 20, 10), seed=10, broken_share=1.0)
    fr = W.frags_of(sc, mats)
    o, d = _both(sc, fr, _seed(sc), mats, ctx)
    assert len(o["members"]) == 1 and len(o["bonds"]) == 0 and o["any_support"] == 0
    _same(o, d)


def test_negative_coordinates(mats, ctx):
    sc = W.scene(shape=(44, 38, 22), seed=11, broken_share=0.02, lo_shift=(-40, -7, -33))
    fr = W.frags_of(sc, mats)
    o, d = _both(sc, fr, _seed(sc), mats, ctx, cluster_nodes=300)
    assert o["cell"] > 0
    _same(o, d)


def test_large_block(mats, ctx):
    sc = W.scene(shape=(96, 80, 40), cell=3, seed=12, broken_share=0.01, damage=True)
    fr = W.frags_of(sc, mats)
    o, d = _both(sc, fr, _seed(sc), mats, ctx)
    assert o["cell"] == 8 and len(o["members"]) > 10000
    _same(o, d)


@pytest.mark.parametrize("fault", ["index", "air", "chunk"])
def test_bad_inputs_are_refused(fault, mats, ctx):
    """The device checks the inputs before it walks (a fragment index in range, on solid unanchored
    voxels, inside the chunk of its first voxel) and refuses them with an error, as the host entries do."""
    from opus_destruct_2_amd.stress import SvxgError
    sc = W.scene(shape=(40, 20, 10), seed=13)
    fr = W.frags_of(sc, mats)
    frag = sc["frag"].copy()
    if fault == "index":
        frag[4, 5, 5] = len(fr) + 3
    elif fault == "air":
        z, y, x = np.argwhere(sc["vox"] == 0)[0]
        frag[z, y, x] = 0
    else:
        # a fragment just below x = 32 given a voxel just above it
        f = int(frag[4, 5, 31])
        frag[4, 5, 32] = f
    with pytest.raises(SvxgError):
        voxels.extract_structure(sc["vox"], frag, sc["lo"], sc["h"], fr, _seed(sc), mats, ctx=ctx)
    # the context goes on working
    o, d = _both(sc, fr, _seed(sc), mats, ctx)
    _same(o, d)
