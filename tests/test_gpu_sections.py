"""Bond sections from voxel faces on the device (svxg_bond_sections) against the oracle
(oracle/oracle.cpp orc_bond_sections): every bond bit for bit - frame, centroid, area, moments,
fibres, springs, strengths - on Voronoi blocks with voids, supports and damaged voxels; and the graph
so built solves on the device as on the CPU."""
import numpy as np
import pytest

import oracle
from oracle import CpuProblem
from opus_destruct_2_amd import stress, voxels
from opus_destruct_2_amd.abi import JUNCTION_SAMPLE, SVXG_MIXED

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,cell,voids,seed", [((16, 14, 12), 2, 0.1, 4), ((30, 24, 20), 3, 0.2, 5),
                                                   ((9, 9, 40), 2, 0.0, 6)])
def test_sections_are_the_oracles(shape, cell, voids, seed, mats, ctx):
    vox, frag, lo = voxels.voronoi_block(*shape, cell=cell, voids=voids, seed=seed)
    rng = np.random.default_rng(seed)
    damage = np.where(rng.random(vox.shape) < 0.05, rng.integers(1, 256, vox.shape), 0).astype(np.uint8)
    nodes, heads, faces = voxels.extract(vox, frag, lo, 0.25, mats)
    dev = voxels.bond_sections(nodes, heads, faces, vox, lo, 0.25, mats, damage=damage, ctx=ctx)
    orc = oracle.bond_sections(nodes, heads, faces, vox, lo, 0.25, mats, damage=damage)
    assert len(dev) > 100 and (dev["b"] < 0).any()
    assert dev.tobytes() == orc.tobytes()


def test_graph_from_voxels_solves(mats, ctx):
    vox, frag, lo = voxels.voronoi_block(20, 18, 14, cell=2, voids=0.1, seed=8)
    nodes, heads, faces = voxels.extract(vox, frag, lo, 0.25, mats)
    bonds = voxels.bond_sections(nodes, heads, faces, vox, lo, 0.25, mats, ctx=ctx)
    f = voxels.gravity(nodes)
    p = stress.StressProblem(nodes, bonds, mats, ctx, options=stress.default_options(precision=SVXG_MIXED))
    assert p.assemble()
    u, r = p.solve(f, rtol=1e-10, maxit=2000)
    assert r["converged"]
    o = CpuProblem(nodes, bonds, mats)
    assert o.assemble()
    uo, ro = o.solve(f, rtol=1e-10, maxit=5000)
    assert ro["converged"]
    assert np.abs(u - uo).max() <= 1e-5 * np.abs(uo).max()
    p.close()


def test_sections_edge_cases(mats, ctx):
    vox, frag, lo = voxels.voronoi_block(6, 6, 6, cell=2, voids=0.0, seed=2)
    nodes, heads, faces = voxels.extract(vox, frag, lo, 0.25, mats)
    assert len(voxels.bond_sections(nodes, heads[:0], faces, vox, lo, 0.25, mats, ctx=ctx)) == 0
    bad = heads.copy()
    bad[0]["nfaces"] = len(faces) + 1
    with pytest.raises(stress.SvxgError):
        voxels.bond_sections(nodes, bad, faces, vox, lo, 0.25, mats, ctx=ctx)


@pytest.mark.parametrize("shape,cell,voids,seed,mixed", [((16, 14, 12), 2, 0.1, 4, False),
                                                         ((30, 24, 20), 3, 0.2, 5, False),
                                                         ((9, 9, 40), 2, 0.0, 6, False),
                                                         ((18, 18, 15), 3, 0.1, 7, True)])
def test_lattice_graph_on_device_is_the_host_extraction(shape, cell, voids, seed, mixed, mats, ctx):
    """svxg_lattice_graph (faces found, sorted and grouped, fragments summed, sections, all on the
    device) gives the records of the host extraction followed by the oracle's sections, bit for bit.
    (mixed: every voxel a random material, so fragments hold many; the commonest one names the node)"""
    vox, frag, lo = voxels.voronoi_block(*shape, cell=cell, voids=voids, seed=seed)
    rng = np.random.default_rng(seed)
    if mixed:
        solid = frag >= 0
        vox[solid] = rng.integers(0, 12, int(solid.sum())) + 1
    damage = np.where(rng.random(vox.shape) < 0.05, rng.integers(1, 256, vox.shape), 0).astype(np.uint8)
    nodes, heads, faces = voxels.extract(vox, frag, lo, 0.25, mats)
    bonds = oracle.bond_sections(nodes, heads, faces, vox, lo, 0.25, mats, damage=damage)
    dn, dh, df, db = voxels.lattice_graph(vox, frag, lo, 0.25, mats, damage=damage, ctx=ctx)
    assert dn.tobytes() == nodes.tobytes()
    assert dh.tobytes() == heads.tobytes()
    assert df.tobytes() == faces.tobytes()
    assert db.tobytes() == bonds.tobytes()


# ---- the general case: placed lattices and junction samples (svxg_bond_sections_general) -------------

from tests.test_sections_oracle import cantilever_scene, general_heads, grids_scene, turned  # noqa: E402


@pytest.mark.parametrize("S", [1, 2, 3, 5, 7])
def test_general_sections_are_the_oracles(S, mats, ctx):
    """Three lattices joined by junctions (pairs across lattices, a junction support, faces within the
    turned ones), damaged voxels: every bond bit for bit the oracle's."""
    grids, frags = grids_scene(seed=S)
    nodes, heads, faces, samples = voxels.extract_grids(grids, frags, mats, S=S)
    assert (heads["nsamples"] > 0).sum() >= 10
    dev = voxels.bond_sections_general(nodes, heads, faces, samples, grids, S, mats, ctx=ctx)
    orc = oracle.bond_sections_general(nodes, heads, faces, samples, grids, S, mats)
    assert dev.tobytes() == orc.tobytes()


def test_general_entry_on_one_unplaced_lattice_is_the_lattice_entry(mats, ctx):
    vox, frag, lo = voxels.voronoi_block(20, 16, 12, cell=2, voids=0.1, seed=12)
    rng = np.random.default_rng(12)
    damage = np.where(rng.random(vox.shape) < 0.05, rng.integers(1, 256, vox.shape), 0).astype(np.uint8)
    nodes, heads, faces = voxels.extract(vox, frag, lo, 0.25, mats)
    one = voxels.bond_sections(nodes, heads, faces, vox, lo, 0.25, mats, damage=damage, ctx=ctx)
    G = voxels.Grid(vox, lo, 0.25, damage=damage)
    gen_ = voxels.bond_sections_general(nodes, general_heads(heads), faces, np.zeros(0, JUNCTION_SAMPLE), [G], 3,
                                        mats, ctx=ctx)
    assert one.tobytes() == gen_.tobytes()
    # (the same lattice marked placed by the identity: the general formulas, the oracle's bits)
    G.identity = False
    dev = voxels.bond_sections_general(nodes, general_heads(heads), faces, np.zeros(0, JUNCTION_SAMPLE), [G], 3,
                                       mats, ctx=ctx)
    orc = oracle.bond_sections_general(nodes, general_heads(heads), faces, np.zeros(0, JUNCTION_SAMPLE), [G], 3, mats)
    assert dev.tobytes() == orc.tobytes()


def test_turned_cantilever_on_the_device(mats, ctx):
    """tests/core/test_grids.cpp:55-80 on the device: the cantilever's sections in an unplaced grid
    and in one turned 30 degrees about z and moved, each solved on the device: utilisations within
    1e-6 of each other and of the oracle's."""
    phis = []
    for R, off in ((None, None), (turned(30.0), (3.2, -1.7, 0.0))):
        G, frag = cantilever_scene(R, off)
        nodes, heads, faces, samples = voxels.extract_grids([G], [frag], mats, S=3)
        b = voxels.bond_sections_general(nodes, heads, faces, samples, [G], 3, mats, ctx=ctx)
        assert b.tobytes() == oracle.bond_sections_general(nodes, heads, faces, samples, [G], 3, mats).tobytes()
        p = stress.StressProblem(nodes, b, mats, ctx, options=stress.default_options())
        assert p.assemble()
        u, r = p.solve(voxels.gravity(nodes), rtol=1e-12, maxit=5000)
        assert r["converged"]
        phis.append(p.bond_checks(None)[0])
        p.close()
    assert phis[0].max() > 0.05
    assert np.allclose(phis[1], phis[0], rtol=1e-6, atol=1e-9)


def test_general_sections_edge_cases(mats, ctx):
    grids, frags = grids_scene(seed=2)
    nodes, heads, faces, samples = voxels.extract_grids(grids, frags, mats, S=3)
    assert len(voxels.bond_sections_general(nodes, heads[:0], faces, samples, grids, 3, mats, ctx=ctx)) == 0
    for field, value in (("face", 6), ("sub", 9), ("vg", 3), ("kind", 2)):
        bad = samples.copy()
        bad[0][field] = value
        with pytest.raises(stress.SvxgError):
            voxels.bond_sections_general(nodes, heads, faces, bad, grids, 3, mats, ctx=ctx)
    bad = heads.copy()
    bad[0]["grid"] = 3
    with pytest.raises(stress.SvxgError):
        voxels.bond_sections_general(nodes, bad, faces, samples, grids, 3, mats, ctx=ctx)
    with pytest.raises(stress.SvxgError):
        voxels.bond_sections_general(nodes, heads, faces, samples, grids, 0, mats, ctx=ctx)


@pytest.mark.parametrize("cell,cubes,mixed", [(8, False, False), (8, True, False), (16, False, True), (16, True, False),
                                              (4, True, True)])
def test_clustered_lattice_graph_is_the_host_extraction(cell, cubes, mixed, mats, ctx):
    """svxg_lattice_graph_clustered (fragments' first voxels and cells, labels propagated over the
    links inside cells, clusters ranked, nodes summed, then the graph over the nodes, all on the
    device) gives extract(cell=...)'s nodes, heads, faces and fragment map and the oracle's sections
    of them, bit for bit."""
    vox, frag, lo = voxels.voronoi_block(48, 40, 30, cell=2, voids=0.1, seed=cell + 3 * cubes)
    rng = np.random.default_rng(cell)
    if mixed:
        solid = frag >= 0
        vox[solid] = rng.integers(0, 12, int(solid.sum())) + 1
    damage = np.where(rng.random(vox.shape) < 0.05, rng.integers(1, 256, vox.shape), 0).astype(np.uint8)
    nodes, heads, faces, fn = voxels.extract(vox, frag, lo, 0.25, mats, cell=cell, cubes=cubes)
    bonds = oracle.bond_sections(nodes, heads, faces, vox, lo, 0.25, mats, damage=damage)
    dn, dh, df, db, dfn = voxels.lattice_graph_clustered(vox, frag, lo, 0.25, mats, cell, cubes, damage=damage, ctx=ctx)
    assert len(dn) < int(frag.max()) + 1
    assert np.array_equal(dfn, fn)
    assert dn.tobytes() == nodes.tobytes()
    assert dh.tobytes() == heads.tobytes()
    assert df.tobytes() == faces.tobytes()
    assert db.tobytes() == bonds.tobytes()


def test_clustered_lattice_graph_edge_cases(mats, ctx):
    vox, frag, lo = voxels.voronoi_block(12, 10, 8, cell=2, voids=0.0, seed=2)
    # cell 0: the unclustered graph, each fragment its own node
    dn, dh, df, db, dfn = voxels.lattice_graph_clustered(vox, frag, lo, 0.25, mats, 0, ctx=ctx)
    n2, h2, f2, b2 = voxels.lattice_graph(vox, frag, lo, 0.25, mats, ctx=ctx)
    assert dn.tobytes() == n2.tobytes() and db.tobytes() == b2.tobytes()
    assert np.array_equal(dfn, np.arange(int(frag.max()) + 1))
    for bad in (3, 64, -1):  # (a cell divides the 32-voxel chunk)
        with pytest.raises(stress.SvxgError):
            voxels.lattice_graph_clustered(vox, frag, lo, 0.25, mats, bad, ctx=ctx)
