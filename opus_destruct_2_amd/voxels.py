"""A structure's graph from a voxel lattice (the step before the stress path): nodes from fragments,
bonds from the voxel faces between them, their sections computed on the device
(svxg_bond_sections).

The reference builds the graph while it walks a structure's fragments (World::Impl::extract and
append_nodes, core/src/world/world.cpp:425-865): a node per fragment (clustered for large
structures), a face list per pair of touching fragments and per (fragment, axis, side) resting on an
anchored voxel, and each list turned into a bond by SecAcc::finish_lattice and section_strengths
(core/src/world/world_internal.hpp:129-249, 374-400). Here one lattice is extracted whole, a node a
fragment: nodes in fragment order, bonds ordered by (node a, pairs before supports, node b / axis
and side), the faces of a bond in the lattice's scan order (x fastest). The order is this module's
(the reference's follows its walk); a bond's section depends only on its faces and their order,
which the device and the oracle share.

Vox codes are the reference's (core/include/svx/world/grid.hpp:28-36): 0 air, material id + 1 in
the low seven bits, 0x80 for an anchored voxel.
"""
import ctypes

import numpy as np

from .abi import BOND, NODE, SECTION_FACE, SECTION_HEAD, VoxelBox

ANCHOR = 0x80


def vox_code(material, anchored=False):
    return (int(material) + 1) | (ANCHOR if anchored else 0)


def voronoi_block(nx, ny, nz, cell=2, voids=0.1, ground=True, seed=1, materials=(0, 1, 3, 7)):
    """A block of nx x ny x nz voxels cut into fragments of about cell^3 voxels (the nearest of
    jittered seeds, one a cell), a share `voids` of the fragments removed (air), a material per
    fragment column drawn from `materials`, on an anchored ground layer (z = -1) when `ground`.
    Returns (vox, frag, lo): arrays of shape (z, y, x) over the box lo .. lo + shape - 1, frag the
    fragment index of each solid unanchored voxel (-1 elsewhere), the fragments that hold voxels
    numbered 0.. in order of their seeds."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    sx, sy, sz = -(-nx // cell), -(-ny // cell), -(-nz // cell)
    gz, gy, gx = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    seeds = (np.stack([gx, gy, gz], -1).reshape(-1, 3) + rng.uniform(0.15, 0.85, (sx * sy * sz, 3))) * cell
    tree = cKDTree(seeds)
    vz, vy, vx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    centres = np.stack([vx, vy, vz], -1).reshape(-1, 3) + 0.5
    _, near = tree.query(centres)
    near = near.reshape(nz, ny, nx)
    keep = rng.random(len(seeds)) >= voids
    col_mat = np.asarray(materials)[rng.integers(0, len(materials), (sy, sx))]
    seed_mat = col_mat[(seeds[:, 1] // cell).astype(int).clip(0, sy - 1), (seeds[:, 0] // cell).astype(int).clip(0, sx - 1)]
    solid = keep[near]
    # renumber the kept fragments that hold voxels 0.. in seed order
    held = np.zeros(len(seeds), bool)
    held[near[solid]] = True
    new = np.full(len(seeds), -1, np.int64)
    new[held] = np.arange(int(held.sum()))
    frag = np.where(solid, new[near], -1).astype(np.int32)
    vox = np.where(solid, seed_mat[near] + 1, 0).astype(np.uint8)
    lo = np.array([0, 0, 0], np.int32)
    if ground:
        g = np.full((1, ny, nx), vox_code(6, anchored=True), np.uint8)  # (bedrock)
        vox = np.concatenate([g, vox], 0)
        frag = np.concatenate([np.full((1, ny, nx), -1, np.int32), frag], 0)
        lo = np.array([0, 0, -1], np.int32)
    return vox, frag, lo


def extract(vox, frag, lo, h, mats):
    """(nodes, heads, faces): a node a fragment (its centre of mass and mass from the materials'
    densities), a section head and face list a bond (pairs of touching fragments; a fragment on an
    anchored voxel, one a side of an axis). vox, frag: (z, y, x) arrays over the box from lo."""
    vox = np.ascontiguousarray(vox, np.uint8)
    frag = np.ascontiguousarray(frag, np.int32)
    nzb, nyb, nxb = frag.shape
    nf = int(frag.max()) + 1 if frag.size else 0
    rho = np.asarray(mats["rho"], np.float64)
    solid = frag >= 0
    zz, yy, xx = np.nonzero(solid)
    f = frag[zz, yy, xx]
    mid = (vox[zz, yy, xx] & 0x7F).astype(np.int64) - 1
    m = rho[mid] * (h * h * h)
    # (a voxel's centre is h times its integer coordinates, core/src/frag/fragments.cpp:315)
    cen = np.stack([xx + lo[0], yy + lo[1], zz + lo[2]], -1) * h
    mass = np.bincount(f, weights=m, minlength=nf)
    com = np.stack([np.bincount(f, weights=m * cen[:, k], minlength=nf) for k in range(3)], -1) / mass[:, None]
    # a fragment's material: its most common one (the lower id on a tie)
    counts = np.zeros((nf, 128), np.int64)
    np.add.at(counts, (f, mid), 1)
    nmat = counts.argmax(1).astype(np.uint8)
    nodes = np.zeros(nf, NODE)
    nodes["c"] = com
    nodes["mass"] = mass
    # faces: for each axis, voxel p and its neighbour q = p + e
    recs = []  # (a, kind, key2, scan, px, py, pz, axis, sign, mb)
    flat = lambda z, y, x: (x + nxb * (y + nyb * z)).astype(np.int64)
    for axis in range(3):
        sl_p = [slice(None)] * 3
        sl_q = [slice(None)] * 3
        ax_np = 2 - axis
        sl_p[ax_np] = slice(0, frag.shape[ax_np] - 1)
        sl_q[ax_np] = slice(1, frag.shape[ax_np])
        fp, fq = frag[tuple(sl_p)], frag[tuple(sl_q)]
        vp, vq = vox[tuple(sl_p)], vox[tuple(sl_q)]
        z, y, x = np.nonzero((fp >= 0) | (fq >= 0))
        a_, b_ = fp[z, y, x], fq[z, y, x]
        va, vb = vp[z, y, x], vq[z, y, x]
        scan = flat(z, y, x)
        px, py, pz = x + lo[0], y + lo[1], z + lo[2]
        # pairs of two fragments
        pr = (a_ >= 0) & (b_ >= 0) & (a_ != b_)
        lo_n, hi_n = np.minimum(a_, b_), np.maximum(a_, b_)
        sign = np.where(a_ < b_, 1, -1)
        recs.append(np.stack([lo_n[pr], np.zeros(pr.sum(), np.int64), hi_n[pr], scan[pr], px[pr], py[pr], pz[pr],
                              np.full(pr.sum(), axis), sign[pr], np.zeros(pr.sum(), np.int64)], -1))
        # a fragment on an anchored voxel above it (side -1) or below it (side +1)
        up = (a_ >= 0) & (b_ < 0) & ((vb & ANCHOR) != 0)
        dn = (b_ >= 0) & (a_ < 0) & ((va & ANCHOR) != 0)
        for sel, node, s_, mbv in ((up, a_, 1, vb), (dn, b_, -1, va)):
            k = int(sel.sum())
            recs.append(np.stack([node[sel], np.ones(k, np.int64), np.full(k, 2 * axis + (1 if s_ > 0 else 0)),
                                  scan[sel], px[sel], py[sel], pz[sel], np.full(k, axis), np.full(k, s_),
                                  ((mbv[sel] & 0x7F).astype(np.int64) - 1)], -1))
    R = np.concatenate(recs, 0) if recs else np.zeros((0, 10), np.int64)
    order = np.lexsort((R[:, 7], R[:, 3], R[:, 2], R[:, 1], R[:, 0]))
    R = R[order]
    key = R[:, :3]
    start = np.flatnonzero(np.r_[True, np.any(key[1:] != key[:-1], 1)]) if len(R) else np.zeros(0, np.int64)
    counts_f = np.diff(np.r_[start, len(R)])
    faces = np.zeros(len(R), SECTION_FACE)
    faces["p"] = R[:, 4:7]
    faces["axis"] = R[:, 7]
    heads = np.zeros(len(start), SECTION_HEAD)
    first = R[start]
    pair = first[:, 1] == 0
    heads["a"] = first[:, 0]
    heads["b"] = np.where(pair, first[:, 2], -1)
    heads["axis"] = np.where(pair, 0, first[:, 2] // 2)
    heads["sign"] = np.where(pair, 1, np.where(first[:, 2] % 2 == 1, 1, -1))
    heads["face0"] = start
    heads["nfaces"] = counts_f
    heads["ma"] = nmat[first[:, 0]]
    heads["mb"] = np.where(pair, nmat[np.where(pair, first[:, 2], 0)], first[:, 9])
    heads["strength_a"] = 1.0
    heads["strength_b"] = np.where(pair, 1.0, 1e9)
    # a pair face's sign: +1 when node a is on its lower side; a support face: the side of a
    faces["sign"] = R[:, 8]
    return nodes, heads, faces


def voxel_box(vox, lo, h, damage=None):
    """The svxg_voxel_box of (z, y, x) arrays (kept alive by the caller)."""
    b = VoxelBox()
    for k in range(3):
        b.lo[k] = int(lo[k])
    nz, ny, nx = vox.shape
    b.dim[0], b.dim[1], b.dim[2] = nx, ny, nz
    b.h = float(h)
    b.vox = vox.ctypes.data
    b.damage = None if damage is None else damage.ctypes.data
    return b


def bond_sections(nodes, heads, faces, vox, lo, h, mats, damage=None, ctx=None):
    """svxg_bond_sections: the bonds (BOND records) of the heads and faces, on the device."""
    from . import stress
    ctx = ctx or stress.default_context()
    vox = np.ascontiguousarray(vox, np.uint8)
    damage = None if damage is None else np.ascontiguousarray(damage, np.uint8)
    nodes = np.ascontiguousarray(nodes, NODE)
    heads = np.ascontiguousarray(heads, SECTION_HEAD)
    faces = np.ascontiguousarray(faces, SECTION_FACE)
    mats = np.ascontiguousarray(mats)
    box = voxel_box(vox, lo, h, damage)
    out = np.zeros(len(heads), BOND)
    stress._check(stress.lib().svxg_bond_sections(ctx.h, nodes.ctypes.data, len(nodes), heads.ctypes.data,
                                                  len(heads), faces.ctypes.data, len(faces), ctypes.byref(box),
                                                  mats.ctypes.data, len(mats), out.ctypes.data), "svxg_bond_sections")
    return out


def lattice_graph(vox, frag, lo, h, mats, damage=None, ctx=None):
    """The whole graph on the device (svxg_lattice_graph): (nodes, heads, faces, bonds), the same
    records as extract() followed by bond_sections(). Two calls: the counts, then the records."""
    from . import stress
    ctx = ctx or stress.default_context()
    vox = np.ascontiguousarray(vox, np.uint8)
    frag = np.ascontiguousarray(frag, np.int32)
    damage = None if damage is None else np.ascontiguousarray(damage, np.uint8)
    mats = np.ascontiguousarray(mats)
    box = voxel_box(vox, lo, h, damage)
    nfrag = int(frag.max()) + 1
    nb, nf = ctypes.c_int32(0), ctypes.c_int32(0)
    L = stress.lib()
    stress._check(L.svxg_lattice_graph(ctx.h, ctypes.byref(box), frag.ctypes.data, nfrag, mats.ctypes.data, len(mats),
                                       None, ctypes.byref(nb), ctypes.byref(nf), None, None, None), "svxg_lattice_graph")
    nodes = np.zeros(nfrag, NODE)
    heads = np.zeros(max(1, nb.value), SECTION_HEAD)
    faces = np.zeros(max(1, nf.value), SECTION_FACE)
    bonds = np.zeros(max(1, nb.value), BOND)
    stress._check(L.svxg_lattice_graph(ctx.h, ctypes.byref(box), frag.ctypes.data, nfrag, mats.ctypes.data, len(mats),
                                       nodes.ctypes.data, ctypes.byref(nb), ctypes.byref(nf), heads.ctypes.data,
                                       faces.ctypes.data, bonds.ctypes.data), "svxg_lattice_graph")
    return nodes, heads[:nb.value], faces[:nf.value], bonds[:nb.value]


def gravity(nodes):
    """f: the nodes' weight (f_z = -9.81 m), 6 per node."""
    f = np.zeros(6 * len(nodes))
    f[2::6] = -9.81 * nodes["mass"]
    return f
