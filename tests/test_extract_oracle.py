"""The oracle's extraction walk (oracle/oracle.cpp orc_extract_structure: World::Impl::extract's walk
and append_nodes, core/src/world/world.cpp:425-620, 739-865) pinned by what the walk's definition
fixes, computed here another way (numpy over the box's faces):

- the identities' mixing function is SplitMix64's output function: its published sequence from
  seed 0;
- the members are the seed's component of unbroken fragment faces (scipy), in breadth-first order
  by first touch: each member's earliest member neighbour (its parent) comes before it, and members
  are ordered by (parent, the parent's first face to it in the parent's chunk scan order);
- with no limits and no clusters, the bonds and their faces are the canonical graph's
  (voxels.extract, pinned by tests/test_sections_oracle.py), pair face signs relative to the bond's
  first node, a pair's first node the one of lesser identity;
- a node limit takes a prefix of the unlimited walk; the radius and fresh owners hold fragments
  fixed: every face from a member to a fragment not taken is a support face;
- clusters are cluster_items over the members' cells (pinned in tests/test_sections_oracle.py);
  faces inside a node drop out, the rest stay;
- outside = 1: every member face on the box's boundary is a support face, and the walk says so;
  a structure with no support says so.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle
from opus_destruct_2_amd import voxels
from opus_destruct_2_amd.abi import extract_options
from tests import walk_scenes as W


def test_mix64_is_splitmix64():
    # SplitMix64 from seed 0 (Steele, Lea and Flood 2014; the sequence its authors publish)
    G = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    assert [oracle.mix64((k * G) & (2 ** 64 - 1)) for k in range(4)] == want


def _walk(sc, fr, seed, mats, **kw):
    return oracle.extract_structure(sc["vox"], sc["frag"], sc["lo"], sc["h"], fr, seed, mats, broken=sc["broken"],
                                    damage=sc["damage"], opts=extract_options(**kw))


def _seed(sc, z=3, y=10, x=10):
    f = int(sc["frag"][z, y, x])
    assert f >= 0
    return f


def _component(T, nfrag, seed):
    g = coo_matrix((np.ones(len(T)), (T[:, 0], T[:, 1])), shape=(nfrag, nfrag))
    _, comp = connected_components(g, directed=False)
    return set(np.flatnonzero(comp == comp[seed]).tolist())


@pytest.mark.parametrize("broken_share", [0.0, 0.08])
def test_members_are_the_component_in_breadth_first_order(broken_share, mats):
    sc = W.scene(broken_share=broken_share, seed=4)
    fr = W.frags_of(sc, mats)
    seed = _seed(sc)
    r = _walk(sc, fr, seed, mats)
    T = W.touches(sc)
    members = r["members"].tolist()
    assert members[0] == seed and len(set(members)) == len(members)
    assert set(members) == _component(T, len(fr), seed)
    touch = W.first_touch(T)
    where = {f: k for k, f in enumerate(members)}
    nbr = {}
    for u, w in touch:
        if u in where and w in where:
            nbr.setdefault(where[w], []).append(where[u])
    keys = []
    for k in range(1, len(members)):
        parent = min(j for j in nbr[k])
        assert parent < k
        keys.append((parent, touch[(members[parent], members[k])]))
    assert all(a < b for a, b in zip(keys, keys[1:]))
    assert r["truncated"] == 0 and r["any_support"] == 1 and r["cell"] == 0


def _bond_table(nodes_frag, heads, faces):
    """{bond key over fragments: sorted faces (p, axis, sign relative to the lesser fragment)}"""
    out = {}
    for hd in heads:
        fa = int(nodes_frag[hd["a"]])
        fs = faces[hd["face0"]:hd["face0"] + hd["nfaces"]]
        if hd["b"] >= 0:
            fb = int(nodes_frag[hd["b"]])
            flip = 1 if fa < fb else -1
            key = (min(fa, fb), max(fa, fb))
            rows = sorted((tuple(f["p"]), int(f["axis"]), int(f["sign"]) * flip) for f in fs)
        else:
            key = (fa, -1, int(hd["axis"]), int(hd["sign"]))
            rows = sorted((tuple(f["p"]), int(f["axis"])) for f in fs)
        assert key not in out
        out[key] = rows
    return out


def test_unlimited_graph_is_the_canonical_graph(mats):
    sc = W.scene(seed=5, voids=0.0)
    fr = W.frags_of(sc, mats)
    r = _walk(sc, fr, _seed(sc), mats)
    assert len(r["members"]) == len(fr) and r["cell"] == 0
    nodes, heads, faces = voxels.extract(sc["vox"], sc["frag"], sc["lo"], sc["h"], mats)
    want = _bond_table(np.arange(len(fr)), heads, faces)
    got = _bond_table(r["members"], r["heads"], r["faces"])
    assert set(got) == set(want)
    supports = [k for k in want if len(k) == 4]
    assert supports and all(got[k] == want[k] for k in want)
    # nodes: a fragment's mass as given, its centre its centre of mass within rounding
    m = r["members"]
    assert np.array_equal(r["nodes"]["mass"], fr["mass"][m])
    assert np.allclose(r["nodes"]["c"], fr["com"][m], rtol=1e-14, atol=1e-14)
    # a pair bond's first node: the lesser identity; the faces' signs as the canonical graph's
    pair = r["heads"]["b"] >= 0
    ia, ib = r["node_ident"][r["heads"]["a"][pair]], r["node_ident"][r["heads"]["b"][pair]]
    assert np.all(ia < ib)
    # the bonds: the sections of their heads and faces
    b = oracle.bond_sections(r["nodes"], r["heads"], r["faces"], sc["vox"], sc["lo"], sc["h"], mats)
    assert b.tobytes() == r["bonds"].tobytes()


def _support_faces(r):
    hs = r["heads"]
    out = []
    for hd in hs[hs["b"] < 0]:
        for f in r["faces"][hd["face0"]:hd["face0"] + hd["nfaces"]]:
            out.append((tuple(int(c) for c in f["p"]), int(f["axis"])))
    return sorted(out)


def _member_faces_to(sc, members_set, pred):
    """Faces (lower voxel, axis) from a member's voxel to a voxel for which pred(vox, frag) holds,
    unbroken."""
    frag, vox, broken, lo = sc["frag"], sc["vox"], sc["broken"], sc["lo"]
    inm = np.isin(frag, np.array(sorted(members_set), np.int32)) & (frag >= 0)
    out = []
    for a in range(3):
        ax = 2 - a
        lo_s = [slice(None)] * 3
        hi_s = [slice(None)] * 3
        lo_s[ax], hi_s[ax] = slice(0, frag.shape[ax] - 1), slice(1, frag.shape[ax])
        ok_b = np.ones(frag[tuple(lo_s)].shape, bool) if broken is None else ((broken[tuple(lo_s)] >> a) & 1) == 0
        for mem_side, other_side in ((lo_s, hi_s), (hi_s, lo_s)):
            sel = inm[tuple(mem_side)] & pred(vox[tuple(other_side)], frag[tuple(other_side)]) & ok_b
            z, y, x = np.nonzero(sel)
            for zz, yy, xx in zip(z, y, x):
                out.append(((int(xx) + int(lo[0]), int(yy) + int(lo[1]), int(zz) + int(lo[2])), a))
    return sorted(out)


def test_node_limit_takes_a_prefix_and_holds_the_rest(mats):
    sc = W.scene(seed=6)
    fr = W.frags_of(sc, mats)
    seed = _seed(sc)
    full = _walk(sc, fr, seed, mats)
    N = len(full["members"]) // 3
    r = _walk(sc, fr, seed, mats, max_nodes=N)
    assert np.array_equal(r["members"], full["members"][:N]) and r["truncated"] == 1
    mem = set(r["members"].tolist())
    want = _member_faces_to(sc, mem, lambda v, f: ((v & 0x80) != 0) | ((f >= 0) & ~np.isin(f, list(mem))))
    assert _support_faces(r) == want


def test_radius_and_owners_hold_fragments(mats):
    sc = W.scene(seed=7)
    fr = W.frags_of(sc, mats)
    seed = _seed(sc)
    R = 2.0
    d = np.linalg.norm(fr["com"] - fr["com"][seed], axis=1)
    # owners: a fresh structure's fragments near the seed are held; an older one's far ones taken
    near = np.flatnonzero((d > 0.6) & (d < 1.2))[:4]
    far = np.flatnonzero((d > R) & (d < R + 0.8))[:25]
    fr["owner"][near] = 9
    fr["owner"][far] = 4
    r = _walk(sc, fr, seed, mats, max_radius=R, fresh_from=8)
    mem = set(r["members"].tolist())
    assert r["truncated"] == 1 and not (mem & set(near.tolist()))
    for f in mem:
        assert d[f] <= R or fr["owner"][f] == 4
    T = W.touches(sc)
    held = {int(w) for u, w, _ in T if u in mem and w not in mem}
    assert held and all(d[w] > R or fr["owner"][w] == 9 for w in held)
    assert set(far.tolist()) & mem  # (owned by a structure of an earlier refresh: taken past the radius)
    want = _member_faces_to(sc, mem, lambda v, f: ((v & 0x80) != 0) | ((f >= 0) & ~np.isin(f, list(mem))))
    assert _support_faces(r) == want


@pytest.mark.parametrize("cubes", [0, 1])
@pytest.mark.parametrize("limit", [200, 700])
def test_clusters_are_cluster_items_of_the_members(cubes, limit, mats):
    sc = W.scene(shape=(48, 40, 24), seed=8, broken_share=0.03)
    fr = W.frags_of(sc, mats)
    r = _walk(sc, fr, _seed(sc), mats, cluster_nodes=limit, cluster_cubes=cubes)
    m = r["members"]
    assert r["cell"] == voxels.cluster_cell(len(m), limit) > 0
    keys = voxels.cell_keys(fr["first"][m], r["cell"], bool(cubes))
    # (a cube key: the cell in the chunk; the default one keeps a chunk's cells apart along x only)
    chunk = np.floor_divide(fr["first"][m], voxels.CHUNK)
    keys = np.concatenate([keys, chunk], 1)
    where = {int(f): k for k, f in enumerate(m)}
    T = W.touches(sc)
    links = np.array([(where[u], where[w]) for u, w, _ in T if u in where and w in where], np.int64)
    cl, nn = voxels.cluster_items(keys, links)
    assert np.array_equal(cl, r["member_node"]) and nn == len(r["nodes"])
    mass = np.zeros(nn)
    for k in range(len(m)):
        mass[cl[k]] += fr["mass"][m[k]]
    assert np.array_equal(mass, r["nodes"]["mass"])
    # faces: every unbroken face between members of two nodes once, none inside a node
    fn = np.full(len(fr), -1)
    fn[m] = cl
    got = sorted((tuple(int(c) for c in f["p"]), int(f["axis"])) for hd in r["heads"][r["heads"]["b"] >= 0]
                 for f in r["faces"][hd["face0"]:hd["face0"] + hd["nfaces"]])
    frag = sc["frag"]
    want = []
    for a in range(3):
        ax = 2 - a
        lo_s, hi_s = [slice(None)] * 3, [slice(None)] * 3
        lo_s[ax], hi_s[ax] = slice(0, frag.shape[ax] - 1), slice(1, frag.shape[ax])
        fa, fb = frag[tuple(lo_s)], frag[tuple(hi_s)]
        na = np.where(fa >= 0, fn[np.maximum(fa, 0)], -1)
        nb = np.where(fb >= 0, fn[np.maximum(fb, 0)], -1)
        ok = (na >= 0) & (nb >= 0) & (na != nb) & (((sc["broken"][tuple(lo_s)] >> a) & 1) == 0)
        z, y, x = np.nonzero(ok)
        want += [((int(xx) + int(sc["lo"][0]), int(yy) + int(sc["lo"][1]), int(zz) + int(sc["lo"][2])), a)
                 for zz, yy, xx in zip(z, y, x)]
    assert got == sorted(want)


def test_outside_holds_what_touches_it(mats):
    # a block cut by its box on the +x side (voxels beyond it unknown) and lifted off the ground
    sc = W.scene(shape=(36, 20, 12), seed=9, voids=0.0)
    sc["vox"] = sc["vox"][1:].copy()
    sc["frag"] = sc["frag"][1:].copy()
    sc["lo"] = sc["lo"] + np.array([0, 0, 1], np.int32)
    fr = W.frags_of(sc, mats)
    seed = _seed(sc, z=2)
    r0 = _walk(sc, fr, seed, mats)
    assert r0["any_support"] == 0 and r0["touched_unloaded"] == 0 and len(_support_faces(r0)) == 0
    r1 = _walk(sc, fr, seed, mats, outside=1)
    assert r1["any_support"] == 1 and r1["touched_unloaded"] == 1
    assert np.array_equal(r1["members"], r0["members"])
    nz, ny, nx = sc["frag"].shape
    lo = sc["lo"]
    want = []
    for a, n in ((0, nx), (1, ny), (2, nz)):
        ax = 2 - a
        for side in (0, n - 1):
            sl = [slice(None)] * 3
            sl[ax] = side
            face = sc["frag"][tuple(sl)] >= 0
            for i, j in zip(*np.nonzero(face)):
                c = [0, 0, 0]
                rest = [k for k in range(3) if k != a]
                # (z, y, x) order of the remaining two axes
                idx = {2 - k: v for k, v in zip(sorted(rest, reverse=True), (i, j))}
                for k in rest:
                    c[k] = idx[2 - k] + int(lo[k])
                c[a] = side + int(lo[a]) - (1 if side == 0 else 0)
                want.append((tuple(c), a))
    assert _support_faces(r1) == sorted(want)
