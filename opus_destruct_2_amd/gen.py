"""Synthetic workloads of the stress solve (wrapper of libsvxgen.so, the deterministic generator).

Configurations (sizes as generated; realised counts are reported by the callers):
  tower   8 x 8 x 20 cubes of 0.5 m on supports, one structural material a storey
  block   64 x 64 x 48 cubes, 15 % voids, only ground-connected cells, materials per column
  pieces  connected random lattice subsets, log-normal sizes (median 200, clamped 16..20000),
          half on supports, half free bodies pinned at the node nearest their centre of mass
  city    buildings (towers on supports) of 5k..150k nodes, log-normal sizes (median 20k), and
          ground-storey collapses that cut the lowest storey's columns under half the footprint
"""
import ctypes
import os

import numpy as np

from .abi import BOND, MATERIAL, NODE

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsvxgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, i32, u64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_double
        L.svxgen_materials.argtypes = [vp]
        L.svxgen_materials.restype = i32
        L.svxgen_tower.argtypes = [i32, i32, i32, f64, u64]
        L.svxgen_tower.restype = vp
        L.svxgen_block.argtypes = [i32, i32, i32, f64, f64, u64]
        L.svxgen_block.restype = vp
        L.svxgen_piece.argtypes = [i32, i32, f64, u64]
        L.svxgen_piece.restype = vp
        L.svxgen_building.argtypes = [i32, i32, i32, f64, u64]
        L.svxgen_building.restype = vp
        L.svxgen_piece_sizes.argtypes = [i32, f64, f64, i32, i32, u64, vp]
        L.svxgen_counts.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.svxgen_copy.argtypes = [vp, vp, vp, vp]
        L.svxgen_free.argtypes = [vp]
        L.svxgen_uniform.argtypes = [u64, ctypes.c_int64, vp]
        L.svxgen_safe_cuts.argtypes = [i32, vp, i32, vp, vp, i32, u64, vp]
        L.svxgen_safe_cuts.restype = i32
        _LIB = L
    return _LIB


class Graph:
    """A fragment graph: nodes (NODE), bonds (BOND), the gravity load f (6 per node)."""

    def __init__(self, nodes, bonds, f, name=""):
        self.nodes, self.bonds, self.f, self.name = nodes, bonds, f, name

    @property
    def n(self):
        return len(self.nodes)

    @property
    def m(self):
        return len(self.bonds)

    @property
    def nfree(self):
        return int(np.count_nonzero(self.nodes["fixed"] == 0))

    @property
    def dofs(self):
        return 6 * self.nfree

    def copy(self):
        return Graph(self.nodes.copy(), self.bonds.copy(), self.f.copy(), self.name)


def _take(h, name):
    L = lib()
    n, m = ctypes.c_int32(), ctypes.c_int32()
    L.svxgen_counts(h, ctypes.byref(n), ctypes.byref(m))
    nodes = np.zeros(n.value, NODE)
    bonds = np.zeros(m.value, BOND)
    f = np.zeros(6 * n.value, np.float64)
    L.svxgen_copy(h, nodes.ctypes.data, bonds.ctypes.data, f.ctypes.data)
    L.svxgen_free(h)
    return Graph(nodes, bonds, f, name)


def materials():
    """The reference's 12 material presets (core/src/material/material.cpp:71-98)."""
    out = np.zeros(12, MATERIAL)
    lib().svxgen_materials(out.ctypes.data)
    return out


def tower(nx=8, ny=8, nz=20, L=0.5, seed=1):
    return _take(lib().svxgen_tower(nx, ny, nz, L, seed), f"tower{nx}x{ny}x{nz}")


def block(nx=64, ny=64, nz=48, L=0.5, voids=0.15, seed=1):
    return _take(lib().svxgen_block(nx, ny, nz, L, voids, seed), f"block{nx}x{ny}x{nz}v{voids}")


def piece(count, pinned=False, L=0.5, seed=1):
    return _take(lib().svxgen_piece(int(count), int(bool(pinned)), L, seed), f"piece{count}{'p' if pinned else 's'}")


def building(nx, ny, nz, L=0.5, seed=1):
    return _take(lib().svxgen_building(nx, ny, nz, L, seed), f"building{nx}x{ny}x{nz}")


def city_shapes(count, median=20000.0, sigma=0.8, lo=5000, hi=150000, seed=1):
    """(nx, ny, nz) of `count` buildings of a city chunk: node counts log-normal (median, sigma)
    clamped to [lo, hi], a footprint of s x s or s x (s+1) cells and a height about r*s storeys,
    r in [1, 3) (taller than wide). Integer arithmetic only past the draws."""
    sizes = piece_sizes(count, median, sigma, lo, hi, seed)
    u = uniform(seed * 7 + 3, 2 * count)
    out = []
    for k, n in enumerate(sizes):
        r = 1.0 + 2.0 * float(u[2 * k])
        s = 1
        while (s + 1) ** 3 * r <= n:
            s += 1
        nx, ny = s, s + (1 if u[2 * k + 1] >= 0.5 else 0)
        out.append((nx, ny, max(3, (int(n) + nx * ny // 2) // (nx * ny))))
    return out


def ground_storey_cuts(g, frac=0.5):
    """A ground-storey collapse: the bonds between the two lowest storeys under the part of the
    footprint with x below `frac` of its width (the support bonds stay). The rest of the footprint
    still carries the storeys above, so no free node loses its path to a support."""
    c = g.nodes["c"]
    a, b = g.bonds["a"], g.bonds["b"]
    _, lay = np.unique(c[:, 2], return_inverse=True)
    x = c[:, 0]
    xcut = x.min() + frac * (x.max() - x.min())
    inner = b >= 0
    bb = np.where(inner, b, 0)
    lo, hi = np.minimum(lay[a], lay[bb]), np.maximum(lay[a], lay[bb])
    pick = inner & (lo == 0) & (hi == 1) & (x[a] < xcut) & (x[bb] < xcut)
    return np.nonzero(pick)[0].astype(np.int32)


def wind_load(g, angle, scale=0.1):
    """g's gravity load plus a horizontal load in direction `angle` (radians) on every node,
    scale * its weight * its height over the building's height (0 at the lowest storey)."""
    c = g.nodes["c"]
    z = c[:, 2] - c[:, 2].min()
    h = max(float(z.max()), 1e-12)
    w = scale * 9.81 * g.nodes["mass"] * (z / h)
    f = g.f.copy()
    f[0::6] += w * np.cos(angle)
    f[1::6] += w * np.sin(angle)
    return f


def safe_cuts(g, broken, count, seed):
    """Up to `count` intact bonds of g whose removal leaves no free node without a path to a support
    or pin (broken: uint8 per bond, 1 = already broken)."""
    broken = np.ascontiguousarray(broken, np.uint8)
    out = np.zeros(max(0, int(count)), np.int32)
    k = lib().svxgen_safe_cuts(g.n, g.nodes.ctypes.data, len(g.bonds), g.bonds.ctypes.data, broken.ctypes.data,
                               int(count), seed, out.ctypes.data)
    return out[:k]


def piece_sizes(count, median=200.0, sigma=0.75, lo=16, hi=20000, seed=1):
    out = np.zeros(count, np.int32)
    lib().svxgen_piece_sizes(count, median, sigma, lo, hi, seed, out.ctypes.data)
    return out


def pieces(count, median=200.0, sigma=0.75, lo=16, hi=20000, seed=1):
    """A batch of `count` pieces; even indices stand on supports, odd ones are pinned."""
    sizes = piece_sizes(count, median, sigma, lo, hi, seed)
    return [piece(int(s), pinned=bool(k & 1), seed=seed * 100003 + k) for k, s in enumerate(sizes)]


def uniform(seed, count):
    out = np.zeros(count, np.float64)
    lib().svxgen_uniform(seed, count, out.ctypes.data)
    return out
