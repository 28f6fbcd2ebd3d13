"""Scenes for the extraction walk's tests: Voronoi blocks cut into chunk-local fragments, with broken
faces, damage, owners; and the numpy facts the walk's output is checked against (the face
adjacency of fragments, each pair's first touch in a fragment's scan order)."""
import numpy as np

from opus_destruct_2_amd import voxels

H = 0.25


def scene(shape=(40, 36, 20), cell=3, voids=0.1, seed=3, broken_share=0.0, damage=False, lo_shift=(0, 0, 0)):
    """A Voronoi block (voxels.voronoi_block) on an anchored ground layer, its fragments cut at chunk
    boundaries, their FRAGMENT records; broken: each voxel face broken with probability
    broken_share (bit a: the face to the +a neighbour)."""
    vox, frag, lo = voxels.voronoi_block(*shape, cell=cell, voids=voids, seed=seed)
    lo = np.asarray(lo, np.int32) + np.asarray(lo_shift, np.int32)
    frag = voxels.chunk_fragments(frag, lo)
    rng = np.random.default_rng(seed + 100)
    broken = None
    if broken_share > 0:
        bits = [(rng.random(vox.shape) < broken_share).astype(np.uint8) << a for a in range(3)]
        broken = (bits[0] | bits[1] | bits[2]).astype(np.uint8)
    dmg = rng.integers(0, 200, vox.shape).astype(np.uint8) if damage else None
    return {"vox": vox, "frag": frag, "lo": lo, "h": H, "broken": broken, "damage": dmg}


def frags_of(sc, mats):
    return voxels.fragments(sc["vox"], sc["frag"], sc["lo"], sc["h"], mats)


def touches(sc):
    """Every unbroken face of a fragment voxel to a voxel of another fragment: (u, w, key) with key =
    the voxel's place in the box's x-slowest scan (t = (x * ny + y) * nz + z) times 6 plus the face
    (-x, +x, -y, +y, -z, +z)."""
    frag, broken = sc["frag"], sc["broken"]
    nz, ny, nx = frag.shape
    out = []
    for a in range(3):
        ax = 2 - a
        for sg in (-1, 1):
            d = 2 * a + (1 if sg > 0 else 0)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if sg > 0:
                src[ax], dst[ax] = slice(0, frag.shape[ax] - 1), slice(1, frag.shape[ax])
            else:
                src[ax], dst[ax] = slice(1, frag.shape[ax]), slice(0, frag.shape[ax] - 1)
            fu, fw = frag[tuple(src)], frag[tuple(dst)]
            ok = (fu >= 0) & (fw >= 0) & (fu != fw)
            if broken is not None:
                lower = broken[tuple(src)] if sg > 0 else broken[tuple(dst)]
                ok &= ((lower >> a) & 1) == 0
            zyx = [c.astype(np.int64) for c in np.nonzero(ok)]
            if sg < 0:
                zyx[ax] += 1  # (the source voxel lies one further along the axis than its slice index)
            zz, yy, xx = zyx
            t = (xx * ny + yy) * nz + zz
            out.append(np.stack([fu[ok], fw[ok], t * 6 + d], -1))
    return np.concatenate(out, 0)


def first_touch(T):
    """{(u, w): the least key of u's faces to w}."""
    order = np.lexsort((T[:, 2], T[:, 1], T[:, 0]))
    S = T[order]
    head = np.r_[True, np.any(S[1:, :2] != S[:-1, :2], 1)]
    return {(int(u), int(w)): int(k) for u, w, k in S[head]}
