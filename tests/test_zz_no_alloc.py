"""The allocation-free bond checks (svxg_problem_bond_checks_noalloc) against svxg_problem_bond_checks:
bit-identical phi and modes over device arrays, from the resident solution and from a u given on the
device, on two problems in a row (the second smaller), and the size and argument errors. Device
arrays come from the HIP runtime the library itself loads (hipMalloc through ctypes)."""
import ctypes

import numpy as np
import pytest

from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_BAD_INPUT, SVXG_MIXED, SVXG_OK
from tests import problems as P

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DevArray:
    def __init__(self, L, nbytes):
        self.L, self.p = L, ctypes.c_void_p()
        assert L.hipMalloc(ctypes.byref(self.p), max(1, nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.L.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.L.hipFree(self.p)


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def check_one(L, prob, g, fragility):
    m, n = len(g.bonds), len(g.nodes)
    phi_ref, mode_ref = prob.bond_checks(None, fragility)
    dphi, dmode = DevArray(L, 8 * m), DevArray(L, m)
    st = stress.lib().svxg_problem_bond_checks_noalloc(prob.h, None, fragility, dphi.p, dmode.p)
    assert st == SVXG_OK, stress.lib().svxg_last_error()
    assert np.array_equal(bits(dphi.get(np.float64, m)), bits(phi_ref))
    assert np.array_equal(dmode.get(np.uint8, m), mode_ref)
    # (a u of the caller's on the device, read in place; the resident solution stays as it was)
    u2 = prob.current() * 1.75
    du = DevArray(L, 8 * 6 * n)
    du.put(u2)
    st = stress.lib().svxg_problem_bond_checks_noalloc(prob.h, du.p, fragility, dphi.p, None)
    assert st == SVXG_OK
    phi2, _ = prob.bond_checks(u2, fragility)
    assert np.array_equal(bits(dphi.get(np.float64, m)), bits(phi2))
    for d in (dphi, dmode, du):
        d.free()


def test_noalloc_bond_checks_two_calls(mats, ctx):
    L = hip()
    big = gen.tower()
    nodes, bonds = P.cantilever(12, 0.5, 300.0)
    small = gen.Graph(nodes, bonds, P.gravity(nodes), "cantilever12")
    pb = stress.StressProblem(big.nodes, big.bonds, mats, ctx, options=stress.default_options(precision=SVXG_MIXED))
    assert pb.assemble()
    pb.solve(big.f, rtol=1e-8)
    ps = stress.StressProblem(nodes, bonds, mats, ctx, options=stress.default_options(amg_min_nodes=4))
    assert ps.assemble()
    ps.solve(small.f, rtol=1e-8)
    check_one(L, pb, big, 1.0)
    check_one(L, ps, small, 2.5)
    pb.close()
    ps.close()


def test_noalloc_bond_checks_errors(mats, ctx):
    nodes, bonds = P.cantilever(6, 0.5, 300.0)
    p = stress.StressProblem(nodes, bonds, mats, ctx, options=stress.default_options(amg_min_nodes=4))
    assert stress.lib().svxg_problem_bond_checks_noalloc(p.h, None, 1.0, None, None) == SVXG_BAD_INPUT
    assert p.assemble()
    L = hip()
    d = DevArray(L, 8 * len(bonds))
    # (assembled but never solved: no resident solution)
    assert stress.lib().svxg_problem_bond_checks_noalloc(p.h, None, 1.0, d.p, None) == SVXG_BAD_INPUT
    d.free()
    p.close()
