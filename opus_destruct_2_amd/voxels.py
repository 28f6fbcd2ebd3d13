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

from .abi import (BOND, FRAGMENT, JUNCTION_SAMPLE, NODE, SECTION_FACE, SECTION_HEAD, SECTION_HEAD_GENERAL, Lattice,
                  VoxelBox)

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


CHUNK = 32  # (the reference's chunk edge, core/include/svx/world/grid.hpp:41)


def cluster_cell(fragments, limit=2500):
    """World::Impl::cluster_cell (world.cpp:733-737): no clusters up to `limit` fragments
    (WorldConfig::cluster_nodes, 2500), cells of 8 voxels up to six times that, else 16."""
    if fragments <= limit:
        return 0
    return 8 if fragments <= 6 * limit else 16


def cell_keys(first, cell, cubes=False):
    """A fragment's cell (append_nodes' key, world.cpp:741-755) from the global coordinates of its
    first voxel (its voxels' least (x, y, z), x first): cubes of `cell` voxels with `cubes`; else
    the reference's default key, whose packing drops the cell along x - cells `cell` voxels in y and
    z and a chunk long in x."""
    first = np.asarray(first, np.int64)
    cx = np.floor_divide(first[:, 0], cell if cubes else CHUNK)
    return np.stack([cx, np.floor_divide(first[:, 1], cell), np.floor_divide(first[:, 2], cell)], -1)


def cluster_items(keys, links):
    """cluster_items (world_internal.hpp:466-491): items with the same key that are linked, directly
    or through each other, form a cluster; clusters numbered in order of first appearance (the least
    item index). keys: (n, k) array; links: (m, 2) item pairs. Returns (cluster of each item, count)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    n = len(keys)
    links = np.asarray(links, np.int64).reshape(-1, 2)
    same = np.all(keys[links[:, 0]] == keys[links[:, 1]], 1) if len(links) else np.zeros(0, bool)
    L = links[same]
    g = coo_matrix((np.ones(len(L)), (L[:, 0], L[:, 1])), shape=(n, n))
    _, comp = connected_components(g, directed=False)
    root = np.full(comp.max() + 1 if n else 0, n, np.int64)
    np.minimum.at(root, comp, np.arange(n))
    order = np.argsort(root, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return rank[comp].astype(np.int32), len(order)


def extract(vox, frag, lo, h, mats, cell=0, cubes=False):
    """(nodes, heads, faces): a node a fragment (its centre of mass and mass from the materials'
    densities), a section head and face list a bond (pairs of touching fragments; a fragment on an
    anchored voxel, one a side of an axis). vox, frag: (z, y, x) arrays over the box from lo.
    cell > 0: touching fragments whose first voxels share a cell (cell_keys) make one node
    (append_nodes, world.cpp:741-800): its mass the fragments' masses summed in index order, its
    centre their centres weighted by their masses (summed, then times 1 / mass), its material the
    heaviest fragment's; faces between fragments of one node drop out. Returns
    (nodes, heads, faces, frag_node) then."""
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
    frag_node = None
    if cell > 0:
        # each fragment's first voxel, its cell, the links between touching fragments, the clusters
        g = np.stack([xx + lo[0], yy + lo[1], zz + lo[2]], -1).astype(np.int64)
        B = 1 << 20
        packed = ((g[:, 0] + B) << 42) | ((g[:, 1] + B) << 21) | (g[:, 2] + B)
        fmin = np.full(nf, np.iinfo(np.int64).max, np.int64)
        np.minimum.at(fmin, f, packed)
        first = np.stack([(fmin >> 42) - B, ((fmin >> 21) & ((1 << 21) - 1)) - B, (fmin & ((1 << 21) - 1)) - B], -1)
        links = []
        for axis in range(3):
            sl_p, sl_q = [slice(None)] * 3, [slice(None)] * 3
            sl_p[2 - axis] = slice(0, frag.shape[2 - axis] - 1)
            sl_q[2 - axis] = slice(1, frag.shape[2 - axis])
            a_, b_ = frag[tuple(sl_p)].ravel(), frag[tuple(sl_q)].ravel()
            pr = (a_ >= 0) & (b_ >= 0) & (a_ != b_)
            links.append(np.stack([a_[pr], b_[pr]], -1))
        frag_node, nn = cluster_items(cell_keys(first, cell, cubes), np.concatenate(links, 0))
        # nodes: the members in index order (bincount sums in input order)
        cmass = np.bincount(frag_node, weights=mass, minlength=nn)
        ccom = np.stack([np.bincount(frag_node, weights=com[:, k] * mass, minlength=nn) for k in range(3)], -1)
        order = np.lexsort((np.arange(nf), -mass, frag_node))
        heavy = order[np.r_[True, frag_node[order][1:] != frag_node[order][:-1]]]
        nmat = nmat[heavy]
        nodes = np.zeros(nn, NODE)
        nodes["c"] = ccom * (1.0 / cmass)[:, None]
        nodes["mass"] = cmass
        frag = np.where(frag >= 0, frag_node[np.maximum(frag, 0)], -1).astype(np.int32)
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
    if cell > 0:
        return nodes, heads, faces, frag_node
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


# ---- several lattices: placed grids and junctions (docs of the reference: GRIDS) ---------------------

class Grid:
    """A voxel lattice placed in the frame (LatticeXf, core/include/svx/phys/rigid.hpp:85-95):
    to(s) = off + R s; R None: unplaced (the identity). vox, damage: (z, y, x) arrays over the box
    from lo, Vox codes."""

    def __init__(self, vox, lo, h, R=None, off=None, damage=None):
        self.vox = np.ascontiguousarray(vox, np.uint8)
        self.lo = np.asarray(lo, np.int32)
        self.h = float(h)
        self.identity = R is None and (off is None or not np.any(off))
        self.R = np.eye(3) if R is None else np.asarray(R, np.float64)
        self.off = np.zeros(3) if off is None else np.asarray(off, np.float64)
        self.damage = None if damage is None else np.ascontiguousarray(damage, np.uint8)

    def to(self, s):
        s = np.asarray(s, np.float64)
        return s if self.identity else s @ self.R.T + self.off

    def dir_to(self, d):
        d = np.asarray(d, np.float64)
        return d if self.identity else d @ self.R.T

    def from_frame(self, x):
        x = np.asarray(x, np.float64)
        return x if self.identity else (x - self.off) @ self.R

    def at(self, p):
        """Vox codes at integer points p (k, 3); air outside the box."""
        p = np.asarray(p, np.int64)
        nz, ny, nx = self.vox.shape
        r = p - self.lo
        inside = (r >= 0).all(1) & (r[:, 0] < nx) & (r[:, 1] < ny) & (r[:, 2] < nz)
        out = np.zeros(len(p), np.uint8)
        out[inside] = self.vox[r[inside, 2], r[inside, 1], r[inside, 0]]
        return out

    def lattice(self):
        L = Lattice()
        for k in range(3):
            L.off[k] = float(self.off[k])
            L.lo[k] = int(self.lo[k])
        for k in range(9):
            L.R[k] = float(self.R.reshape(9)[k])
        L.h = self.h
        L.identity = int(bool(self.identity))
        nz, ny, nx = self.vox.shape
        L.dim[0], L.dim[1], L.dim[2] = nx, ny, nz
        L.vox = self.vox.ctypes.data
        L.damage = None if self.damage is None else self.damage.ctypes.data
        return L


def rotation(axis, deg):
    """The rotation matrix of `deg` degrees about `axis` (Rodrigues)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = np.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * (K @ K)


def lattices(grids):
    arr = (Lattice * len(grids))()
    for g, G in enumerate(grids):
        arr[g] = G.lattice()
    return arr


def junction_points(v, face, sub, S, h, push=0.0):
    """junction_point (world_internal.hpp:94-103) of many samples: (k, 3) lattice points (metres)."""
    v = np.asarray(v, np.float64)
    face = np.asarray(face, np.int64)
    sub = np.asarray(sub, np.int64)
    a = face >> 1
    a1, a2 = (a + 1) % 3, (a + 2) % 3
    sg = np.where(face & 1, 1.0, -1.0)
    i, j = sub // S, sub % S
    s = h * v
    rows = np.arange(len(s))
    s[rows, a] += sg * (0.5 * h + push)
    s[rows, a1] += ((i + 0.5) / S - 0.5) * h
    s[rows, a2] += ((j + 0.5) / S - 0.5) * h
    return s


def extract_grids(grids, frags, mats, S=3, push=None):
    """(nodes, heads, faces, samples) of a structure over several lattices, the multi-lattice analogue
    of extract(): frags[g] holds each voxel's node (global over the grids; -1 for none). Within a
    lattice, faces as extract() finds them; between lattices, junction samples: each face of a node's
    voxel that looks into air of its own lattice is split S x S, each sample point (pushed out of the
    face by `push`, h / 8 by default) is looked up in the older lattices (lower index first from the
    newest down) and, on a solid voxel, joins the two nodes (or, on an anchored voxel, rests the node
    on it: a junction support). Bonds ordered by (node a, pairs before supports, node b / key), faces
    and samples in scan order. Node centres and masses in the frame."""
    rho = np.asarray(mats["rho"], np.float64)
    nn = max(int(f.max()) + 1 if f.size else 0 for f in frags)
    mass = np.zeros(nn)
    mom = np.zeros((nn, 3))
    counts = np.zeros((nn, 128), np.int64)
    frec, srec = [], []  # face: (a, kind, key2, g, scan, px, py, pz, axis, sign, mb); sample: (a, kind, key2, g, scan, vx, vy, vz, ox, oy, oz, vg, og, face, sub, kind, sign)
    for g, (G, frag) in enumerate(zip(grids, frags)):
        frag = np.ascontiguousarray(frag, np.int32)
        vox = G.vox
        nzb, nyb, nxb = frag.shape
        zz, yy, xx = np.nonzero(frag >= 0)
        f = frag[zz, yy, xx]
        mid = (vox[zz, yy, xx] & 0x7F).astype(np.int64) - 1
        m = rho[mid] * G.h ** 3
        cen = G.to(np.stack([xx + G.lo[0], yy + G.lo[1], zz + G.lo[2]], -1) * G.h)
        np.add.at(mass, f, m)
        np.add.at(mom, f, m[:, None] * cen)
        np.add.at(counts, (f, mid), 1)
        flat = lambda z, y, x: (x + nxb * (y + nyb * z)).astype(np.int64)
        for axis in range(3):
            sl_p, sl_q = [slice(None)] * 3, [slice(None)] * 3
            ax_np = 2 - axis
            sl_p[ax_np] = slice(0, frag.shape[ax_np] - 1)
            sl_q[ax_np] = slice(1, frag.shape[ax_np])
            fp, fq = frag[tuple(sl_p)], frag[tuple(sl_q)]
            vp, vq = vox[tuple(sl_p)], vox[tuple(sl_q)]
            z, y, x = np.nonzero((fp >= 0) | (fq >= 0))
            a_, b_ = fp[z, y, x], fq[z, y, x]
            va, vb = vp[z, y, x], vq[z, y, x]
            scan = flat(z, y, x)
            px, py, pz = x + G.lo[0], y + G.lo[1], z + G.lo[2]
            pr = (a_ >= 0) & (b_ >= 0) & (a_ != b_)
            k = int(pr.sum())
            frec.append(np.stack([np.minimum(a_, b_)[pr], np.zeros(k, np.int64), np.maximum(a_, b_)[pr],
                                  np.full(k, g), scan[pr], px[pr], py[pr], pz[pr], np.full(k, axis),
                                  np.where(a_ < b_, 1, -1)[pr], np.zeros(k, np.int64)], -1))
            up = (a_ >= 0) & (b_ < 0) & ((vb & ANCHOR) != 0)
            dn = (b_ >= 0) & (a_ < 0) & ((va & ANCHOR) != 0)
            for sel, node, s_, mbv in ((up, a_, 1, vb), (dn, b_, -1, va)):
                k = int(sel.sum())
                frec.append(np.stack([node[sel], np.ones(k, np.int64), np.full(k, 2 * axis + (1 if s_ > 0 else 0)),
                                      np.full(k, g), scan[sel], px[sel], py[sel], pz[sel], np.full(k, axis),
                                      np.full(k, s_), ((mbv[sel] & 0x7F).astype(np.int64) - 1)], -1))
        if g == 0:
            continue
        # junction samples of this lattice's open faces into the older lattices
        pz_ = push if push is not None else G.h / 8
        vv = np.stack([xx + G.lo[0], yy + G.lo[1], zz + G.lo[2]], -1)
        for face in range(6):
            e = np.zeros(3, np.int64)
            e[face >> 1] = 1 if face & 1 else -1
            open_ = G.at(vv + e) == 0
            if not open_.any():
                continue
            vs, fs = vv[open_], f[open_]
            for sub in range(S * S):
                pts = G.to(junction_points(vs, np.full(len(vs), face), np.full(len(vs), sub), S, G.h, pz_))
                found = np.zeros(len(vs), bool)
                for og in range(g - 1, -1, -1):
                    O = grids[og]
                    o = np.floor(O.from_frame(pts) / O.h + 0.5).astype(np.int64)
                    ov = O.at(o)
                    hit = ~found & (ov != 0)
                    if not hit.any():
                        continue
                    found |= hit
                    of = np.full(len(vs), -1, np.int64)
                    r = o[hit] - O.lo
                    of[hit] = frags[og][r[:, 2], r[:, 1], r[:, 0]]
                    anch = hit & ((ov & ANCHOR) != 0)
                    pair = hit & ~anch & (of >= 0) & (of != fs)
                    for sel, kind in ((pair, 0), (anch, 1)):
                        k = int(sel.sum())
                        if not k:
                            continue
                        a = np.where(kind == 0, np.minimum(fs[sel], of[sel]), fs[sel])
                        key2 = np.maximum(fs[sel], of[sel]) if kind == 0 else np.full(k, 6 + og)
                        sign = np.where(a == fs[sel], 1, -1)
                        scan = (np.flatnonzero(open_)[sel] * 6 + face) * (S * S) + sub
                        srec.append(np.stack([a, np.full(k, kind), key2, np.full(k, g), scan,
                                              vs[sel, 0], vs[sel, 1], vs[sel, 2], o[sel, 0], o[sel, 1], o[sel, 2],
                                              np.full(k, g), np.full(k, og), np.full(k, face), np.full(k, sub),
                                              np.zeros(k, np.int64), sign,
                                              ((ov[sel] & 0x7F).astype(np.int64) - 1)], -1))
    nodes = np.zeros(nn, NODE)
    nodes["c"] = mom / np.maximum(mass, 1e-300)[:, None]
    nodes["mass"] = mass
    nmat = counts.argmax(1).astype(np.uint8)
    F = np.concatenate(frec, 0) if frec else np.zeros((0, 11), np.int64)
    J = np.concatenate(srec, 0) if srec else np.zeros((0, 18), np.int64)
    F = F[np.lexsort((F[:, 8], F[:, 4], F[:, 3], F[:, 2], F[:, 1], F[:, 0]))]
    J = J[np.lexsort((J[:, 4], J[:, 3], J[:, 2], J[:, 1], J[:, 0]))]
    keys = sorted(set(map(tuple, F[:, :3].tolist())) | set(map(tuple, J[:, :3].tolist())))
    fk = {k: [] for k in keys}
    for i, r in enumerate(F[:, :3].tolist()):
        fk[tuple(r)].append(i)
    jk = {k: [] for k in keys}
    for i, r in enumerate(J[:, :3].tolist()):
        jk[tuple(r)].append(i)
    faces = np.zeros(len(F), SECTION_FACE)
    samples = np.zeros(len(J), JUNCTION_SAMPLE)
    heads = np.zeros(len(keys), SECTION_HEAD_GENERAL)
    f0 = s0 = 0
    for h_, key in enumerate(keys):
        a, kind, key2 = key
        fi, ji = fk[key], jk[key]
        for q, i in enumerate(fi):
            faces[f0 + q] = ((F[i, 5], F[i, 6], F[i, 7]), F[i, 8], F[i, 9], (0, 0))
        for q, i in enumerate(ji):
            samples[s0 + q] = ((J[i, 5], J[i, 6], J[i, 7]), (J[i, 8], J[i, 9], J[i, 10]), J[i, 11], J[i, 12],
                               J[i, 13], J[i, 14], J[i, 15], J[i, 16], 1.0)
        grid = int(F[fi[0], 3]) if fi else int(J[ji[0], 3])
        pair = kind == 0
        if pair:
            axis, sign, mb = 0, 1, int(nmat[key2])
        elif key2 < 6:
            axis, sign, mb = key2 // 2, (1 if key2 % 2 else -1), int(F[fi[0], 10])
        else:
            axis, sign, mb = 3 + 6 * (key2 - 6), 1, int(J[ji[0], 17])
        heads[h_] = (a, key2 if pair else -1, axis, sign, f0, len(fi), s0, len(ji), grid, int(nmat[a]), mb,
                     (0,) * 4, 1.0, 1.0 if pair else 1e9)
        f0 += len(fi)
        s0 += len(ji)
    return nodes, heads, faces, samples


def bond_sections_general(nodes, heads, faces, samples, grids, S, mats, ctx=None):
    """svxg_bond_sections_general: the bonds (BOND records) of the heads over placed lattices and
    junction samples, on the device."""
    from . import stress
    ctx = ctx or stress.default_context()
    nodes = np.ascontiguousarray(nodes, NODE)
    heads = np.ascontiguousarray(heads, SECTION_HEAD_GENERAL)
    faces = np.ascontiguousarray(faces, SECTION_FACE)
    samples = np.ascontiguousarray(samples, JUNCTION_SAMPLE)
    mats = np.ascontiguousarray(mats)
    lat = lattices(grids)
    out = np.zeros(len(heads), BOND)
    stress._check(stress.lib().svxg_bond_sections_general(
        ctx.h, nodes.ctypes.data, len(nodes), heads.ctypes.data, len(heads), faces.ctypes.data, len(faces),
        samples.ctypes.data, len(samples), lat, len(grids), int(S), mats.ctypes.data, len(mats), out.ctypes.data),
        "svxg_bond_sections_general")
    return out


def lattice_graph_clustered(vox, frag, lo, h, mats, cell, cubes=False, damage=None, ctx=None):
    """svxg_lattice_graph_clustered: (nodes, heads, faces, bonds, frag_node), the records of
    extract(cell=...) followed by bond_sections(), all on the device."""
    from . import stress
    ctx = ctx or stress.default_context()
    vox = np.ascontiguousarray(vox, np.uint8)
    frag = np.ascontiguousarray(frag, np.int32)
    damage = None if damage is None else np.ascontiguousarray(damage, np.uint8)
    mats = np.ascontiguousarray(mats)
    box = voxel_box(vox, lo, h, damage)
    nfrag = int(frag.max()) + 1
    nn, nb, nf = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    fnode = np.zeros(nfrag, np.int32)
    L = stress.lib()
    stress._check(L.svxg_lattice_graph_clustered(ctx.h, ctypes.byref(box), frag.ctypes.data, nfrag, int(cell),
                                                 int(bool(cubes)), mats.ctypes.data, len(mats), fnode.ctypes.data,
                                                 ctypes.byref(nn), None, ctypes.byref(nb), ctypes.byref(nf), None,
                                                 None, None), "svxg_lattice_graph_clustered")
    nodes = np.zeros(max(1, nn.value), NODE)
    heads = np.zeros(max(1, nb.value), SECTION_HEAD)
    faces = np.zeros(max(1, nf.value), SECTION_FACE)
    bonds = np.zeros(max(1, nb.value), BOND)
    stress._check(L.svxg_lattice_graph_clustered(ctx.h, ctypes.byref(box), frag.ctypes.data, nfrag, int(cell),
                                                 int(bool(cubes)), mats.ctypes.data, len(mats), fnode.ctypes.data,
                                                 ctypes.byref(nn), nodes.ctypes.data, ctypes.byref(nb),
                                                 ctypes.byref(nf), heads.ctypes.data, faces.ctypes.data,
                                                 bonds.ctypes.data), "svxg_lattice_graph_clustered")
    return nodes[:nn.value], heads[:nb.value], faces[:nf.value], bonds[:nb.value], fnode


def chunk_fragments(frag, lo):
    """The fragments cut at chunk boundaries (the reference fragments each chunk on its own,
    fragments.cpp:283-345: a fragment lies in one chunk), numbered in order of (fragment, chunk)."""
    frag = np.ascontiguousarray(frag, np.int32)
    nz, ny, nx = frag.shape
    z, y, x = np.nonzero(frag >= 0)
    cx = np.floor_divide(x + int(lo[0]), CHUNK)
    cy = np.floor_divide(y + int(lo[1]), CHUNK)
    cz = np.floor_divide(z + int(lo[2]), CHUNK)
    key = np.stack([frag[z, y, x].astype(np.int64), cx, cy, cz], -1)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    out = np.full(frag.shape, -1, np.int32)
    out[z, y, x] = inv.reshape(-1).astype(np.int32)
    return out


def fragments(vox, frag, lo, h, mats):
    """FRAGMENT records (FragInfo as the walk reads it, fragments.cpp:283-345) of a box's fragments:
    mass = the voxels' densities times h^3, centre of mass (voxel centres at h times their integer
    coordinates), first = the fragment's least voxel in its chunk's scan order (x slowest) and its
    material; strength 1 (undesigned), owner 0 (none)."""
    vox = np.ascontiguousarray(vox, np.uint8)
    frag = np.ascontiguousarray(frag, np.int32)
    nf = int(frag.max()) + 1 if frag.size else 0
    rho = np.asarray(mats["rho"], np.float64)
    z, y, x = np.nonzero(frag >= 0)
    f = frag[z, y, x]
    mid = (vox[z, y, x] & 0x7F).astype(np.int64) - 1
    m = rho[mid] * (h * h * h)
    g = np.stack([x + int(lo[0]), y + int(lo[1]), z + int(lo[2])], -1).astype(np.int64)
    out = np.zeros(nf, FRAGMENT)
    out["mass"] = np.bincount(f, weights=m, minlength=nf)
    for k in range(3):
        out["com"][:, k] = np.bincount(f, weights=m * g[:, k] * h, minlength=nf) / out["mass"]
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], f))
    head = order[np.r_[True, f[order][1:] != f[order][:-1]]]
    out["first"][f[head]] = g[head]
    out["mat"][f[head]] = mid[head]
    out["strength"] = 1.0
    return out


def _extract(call, vox, frag, lo, h, frags, seed, mats, broken, damage, opts):
    """svxg_extract_structure (or its restatement) with room for a structure of every fragment: the
    members and nodes at most nfrag, the faces at most six a fragment voxel, the bonds at most the
    faces (the arrays are zero pages until written); a second call with the exact counts when the
    bonds outgrow the room given. call(box, broken, frag, frags, nfrag, seed, opts, mats, nmats, res,
    *outputs) -> status (0 ok, 1 bad input or short room)."""
    from .abi import ExtractResult, extract_options
    vox = np.ascontiguousarray(vox, np.uint8)
    frag = np.ascontiguousarray(frag, np.int32)
    frags = np.ascontiguousarray(frags, FRAGMENT)
    broken = None if broken is None else np.ascontiguousarray(broken, np.uint8)
    damage = None if damage is None else np.ascontiguousarray(damage, np.uint8)
    mats = np.ascontiguousarray(mats)
    opts = opts if opts is not None else extract_options()
    box = voxel_box(vox, lo, h, damage)
    bp = None if broken is None else broken.ctypes.data
    nsolid = int(np.count_nonzero(frag >= 0))
    room = {"members": len(frags), "nodes": len(frags), "faces": max(1, 6 * nsolid)}
    room["bonds"] = min(room["faces"], 12 * len(frags))
    for attempt in range(2):
        res = ExtractResult(**room)
        out = {"members": np.zeros(room["members"], np.int32), "member_node": np.zeros(room["members"], np.int32),
               "nodes": np.zeros(room["nodes"], NODE), "node_ident": np.zeros(room["nodes"], np.uint64),
               "heads": np.zeros(room["bonds"], SECTION_HEAD), "faces": np.zeros(room["faces"], SECTION_FACE),
               "bonds": np.zeros(room["bonds"], BOND), "bond_ident": np.zeros(room["bonds"], np.uint64)}
        st = call(ctypes.byref(box), bp, frag.ctypes.data, frags.ctypes.data, len(frags), int(seed),
                  ctypes.byref(opts), mats.ctypes.data, len(mats), ctypes.byref(res),
                  *(out[k].ctypes.data for k in ("members", "member_node", "nodes", "node_ident", "heads", "faces",
                                                 "bonds", "bond_ident")))
        counts = {k: getattr(res, k) for k in ("members", "nodes", "bonds", "faces")}
        if st == 0:
            break
        if attempt or all(counts[k] <= room[k] for k in room):
            raise RuntimeError(f"svxg_extract_structure: status {st}")
        room = {k: max(1, counts[k]) for k in room}
    n = {"members": counts["members"], "member_node": counts["members"], "nodes": counts["nodes"],
         "node_ident": counts["nodes"], "heads": counts["bonds"], "faces": counts["faces"], "bonds": counts["bonds"],
         "bond_ident": counts["bonds"]}
    out = {k: v[:n[k]] for k, v in out.items()}  # (views: the room past them was never written)
    for k in ("cell", "truncated", "any_support", "touched_unloaded"):
        out[k] = int(getattr(res, k))
    return out


def extract_structure(vox, frag, lo, h, frags, seed, mats, broken=None, damage=None, opts=None, ctx=None):
    """svxg_extract_structure: the structure the walk from fragment `seed` takes (World::Impl::extract
    and append_nodes, world.cpp:425-865, for the world grid), on the device. vox, frag, broken,
    damage: (z, y, x) arrays over the box from lo; frags: FRAGMENT records (fragments()); opts:
    abi.extract_options(...). Returns a dict: members, member_node, nodes, node_ident, heads, faces,
    bonds, bond_ident, and the walk's cell, truncated, any_support and touched_unloaded."""
    from . import stress
    ctx = ctx or stress.default_context()
    L = stress.lib()

    def call(*a):
        st = L.svxg_extract_structure(ctx.h, *a)
        if st not in (0, 1):
            stress._check(st, "svxg_extract_structure")
        return st
    try:
        return _extract(call, vox, frag, lo, h, frags, seed, mats, broken, damage, opts)
    except RuntimeError as e:
        raise stress.SvxgError(f"{e}: {L.svxg_last_error().decode(errors='replace')}") from None


def gravity(nodes):
    """f: the nodes' weight (f_z = -9.81 m), 6 per node."""
    f = np.zeros(6 * len(nodes))
    f[2::6] = -9.81 * nodes["mass"]
    return f
