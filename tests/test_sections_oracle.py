"""The oracle's bond sections (oracle/oracle.cpp orc_bond_sections: SecAcc::finish_lattice and
section_strengths, core/src/world/world_internal.hpp:129-249, 374-400) pinned by what geometry fixes:

- a single face: area h^2, centroid at the face centre (a voxel's centre is h times its integer
  coordinates), s = h^4 / 12, fibres h / 2, springs L / 2;
- a k x k square interface: area (k h)^2 and s = (k h)^4 / 12 (the parallel-axis sum of the faces'
  own h^4 / 12);
- an L-shaped interface: the frame turned to the section's principal axes - its product of inertia
  vanishes, its moments are the eigenvalues of the inertia tensor (numpy.linalg.eigh), t1 carries the
  larger;
- supports: the normal along the face axis and side, la = max(h / 4, the centroid's depth), lb = h / 2;
- strengths: a face between two materials takes the weaker of each strength, an air side the other
  side's material, damage scales by the condition (1 - damage / 255, at least 1e-6), the section the
  faces' mean;
- a Voronoi block's graph from voxels (opus_destruct_2_amd/voxels.py) assembles and solves.
"""
import numpy as np
import pytest

import oracle
from opus_destruct_2_amd import gen, voxels
from opus_destruct_2_amd.abi import NODE, SECTION_FACE, SECTION_HEAD

H = 0.25


def _box(shape=(6, 6, 6), lo=(-1, -1, -1), mat=1):
    vox = np.full(shape, mat + 1, np.uint8)
    return vox, np.array(lo, np.int32)


def _nodes(*cs):
    n = np.zeros(len(cs), NODE)
    n["c"] = cs
    n["mass"] = 1.0
    return n


def _pair(faces_p, axis, sign=1, a=0, b=1, ma=1, mb=1):
    heads = np.zeros(1, SECTION_HEAD)
    heads[0] = (a, b, 0, 1, 0, len(faces_p), ma, mb, [0] * 6, 1.0, 1.0)
    faces = np.zeros(len(faces_p), SECTION_FACE)
    faces["p"] = faces_p
    faces["axis"] = axis
    faces["sign"] = sign
    return heads, faces


def test_single_face(mats):
    vox, lo = _box()
    nodes = _nodes([0.0, 0.0, 0.0], [H, 0.0, 0.0])
    heads, faces = _pair([[0, 0, 0]], axis=0)
    b = oracle.bond_sections(nodes, heads, faces, vox, lo, H, mats)[0]
    assert np.allclose(b["n"], [1, 0, 0]) and abs(np.dot(b["t1"], b["t2"])) < 1e-15
    assert np.isclose(b["area"], H * H) and np.allclose(b["p"], [0.5 * H, 0.0, 0.0])
    assert np.isclose(b["s1"], H ** 4 / 12) and np.isclose(b["s2"], H ** 4 / 12)
    assert np.isclose(b["c1"], 0.5 * H) and np.isclose(b["rmax"], H / np.sqrt(2))
    assert np.isclose(b["la"], 0.5 * H) and np.isclose(b["lb"], 0.5 * H)
    assert b["faces"] == 1 and b["sectioned"] == 1 and b["tag"] == 0 and b["strength"] == 1.0


@pytest.mark.parametrize("k", [2, 3, 5])
def test_square_interface(k, mats):
    vox, lo = _box((8, 8, 8))
    c = (k - 1) * H / 2
    nodes = _nodes([c, c, 0.0], [c, c, H])
    pts = [[i, j, 0] for j in range(k) for i in range(k)]
    heads, faces = _pair(pts, axis=2)
    b = oracle.bond_sections(nodes, heads, faces, vox, lo, H, mats)[0]
    assert np.allclose(b["n"], [0, 0, 1])
    assert np.isclose(b["area"], (k * H) ** 2)
    assert np.allclose(b["p"], [c, c, 0.5 * H])
    assert np.isclose(b["s1"], (k * H) ** 4 / 12, rtol=1e-12) and np.isclose(b["s2"], (k * H) ** 4 / 12, rtol=1e-12)
    assert np.isclose(b["c1"], k * H / 2) and np.isclose(b["c2"], k * H / 2)


def test_l_shaped_interface_principal_axes(mats):
    vox, lo = _box((8, 8, 8))
    pts = [[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0], [0, 1, 0], [0, 2, 0]]
    nodes = _nodes([0.2, 0.1, 0.0], [0.2, 0.1, H])
    heads, faces = _pair(pts, axis=2)
    b = oracle.bond_sections(nodes, heads, faces, vox, lo, H, mats)[0]
    x = (np.array(pts, float) + [0.0, 0.0, 0.5]) * H
    w = np.full(len(pts), H * H)
    pc = (w[:, None] * x).sum(0) / w.sum()
    assert np.allclose(b["p"], pc)
    d = x - pc
    I = np.einsum("k,ki,kj->ij", w, d[:, :2], d[:, :2])
    lam = np.linalg.eigh(I)[0]
    y1, y2 = d @ b["t1"], d @ b["t2"]
    own = b["area"] * H * H / 12
    assert abs((w * y1 * y2).sum()) <= 1e-12 * lam.sum()  # (no product of inertia left)
    assert np.isclose(b["s1"] - own, lam[1]) and np.isclose(b["s2"] - own, lam[0])
    assert np.isclose(abs(np.dot(np.cross(b["t1"], b["t2"]), b["n"])), 1.0)


def test_support(mats):
    vox, lo = _box()
    nodes = _nodes([0.0, 0.0, 0.0])
    heads = np.zeros(1, SECTION_HEAD)
    heads[0] = (0, -1, 2, -1, 0, 1, 1, 6, [0] * 6, 1.0, 1e9)
    faces = np.zeros(1, SECTION_FACE)
    faces["p"] = [[0, 0, -1]]
    faces["axis"] = 2
    faces["sign"] = -1
    b = oracle.bond_sections(nodes, heads, faces, vox, lo, H, mats)[0]
    assert np.allclose(b["n"], [0, 0, -1]) and b["b"] == -1
    assert np.allclose(b["p"], [0.0, 0.0, -0.5 * H])
    assert np.isclose(b["la"], 0.5 * H) and np.isclose(b["lb"], 0.5 * H)
    assert b["strength"] == 1.0


def test_section_strengths(mats):
    # faces between voxels of materials 1 and 3, one with an air side, one damaged
    vox = np.zeros((4, 4, 4), np.uint8)
    lo = np.array([0, 0, 0], np.int32)
    vox[0, 0, 0] = 2  # material 1 at (0,0,0)
    vox[0, 0, 1] = 4  # material 3 at (1,0,0)
    vox[0, 1, 0] = 2  # material 1 at (0,1,0); (1,1,0) is air
    vox[0, 2, 0] = 2  # material 1 at (0,2,0)
    vox[0, 2, 1] = 2  # material 1 at (1,2,0)
    dmg = np.zeros_like(vox)
    dmg[0, 2, 1] = 51  # (condition 0.8)
    nodes = _nodes([0.0, H, 0.0], [H, H, 0.0])
    heads, faces = _pair([[0, 0, 0], [0, 1, 0], [0, 2, 0]], axis=0)
    b = oracle.bond_sections(nodes, heads, faces, vox, lo, H, mats, damage=dmg)[0]
    m1, m3 = mats[1], mats[3]
    for key, mk in (("ft", "ft"), ("fb", "fb"), ("fc", "fc"), ("coh", "cohesion")):
        want = (min(m1[mk], m3[mk]) + m1[mk] + 0.8 * m1[mk]) / 3
        assert np.isclose(b[key], want, rtol=1e-6), key
    assert np.isclose(b["mu"], (min(m1["friction"], m3["friction"]) + 2 * m1["friction"]) / 3, rtol=1e-6)
    assert b["sectioned"] == 1


def test_voronoi_graph_from_voxels_solves(mats):
    vox, frag, lo = voxels.voronoi_block(16, 14, 12, cell=2, voids=0.1, seed=4)
    nodes, heads, faces = voxels.extract(vox, frag, lo, H, mats)
    assert len(nodes) == frag.max() + 1 and (heads["b"] < 0).any() and np.all(nodes["mass"] > 0)
    assert np.all(heads["nfaces"] > 0) and heads["nfaces"].sum() == len(faces)
    assert np.all((heads["b"] < 0) | (heads["a"] < heads["b"]))
    b = oracle.bond_sections(nodes, heads, faces, vox, lo, H, mats)
    assert np.all(b["area"] > 0) and np.all(b["la"] > 0) and np.all(b["lb"] > 0)
    assert np.allclose(np.einsum("ij,ij->i", b["n"], b["t1"]), 0, atol=1e-12)
    p = oracle.CpuProblem(nodes, b, mats)
    assert p.assemble()
    u, r = p.solve(voxels.gravity(nodes), rtol=1e-10, maxit=5000)
    assert r["converged"]
    # (the weight is carried: the support bonds' forces, N n + V1 t1 + V2 t2, add up to it)
    loads = p.bond_loads(u)
    sup = b["b"] < 0
    Fz = loads["N"][sup] * b["n"][sup, 2] + loads["V1"][sup] * b["t1"][sup, 2] + loads["V2"][sup] * b["t2"][sup, 2]
    assert np.isclose(Fz.sum(), 9.81 * nodes["mass"].sum(), rtol=1e-6)
