"""Bond sections from voxel faces on the device (svxg_bond_sections) against the oracle
(oracle/oracle.cpp orc_bond_sections): every bond bit for bit - frame, centroid, area, moments,
fibres, springs, strengths - on Voronoi blocks with voids, supports and damaged voxels; and the graph
so built solves on the device as on the CPU."""
import numpy as np
import pytest

import oracle
from oracle import CpuProblem
from opus_destruct_2_amd import stress, voxels
from opus_destruct_2_amd.abi import SVXG_MIXED

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
