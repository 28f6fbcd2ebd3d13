"""The reference's own stress tests (tests/core/test_stress.cpp) restated on the device library.

Closed forms (cantilever statics, chain tension, support reactions), agreement of
preconditioners, in-place updates equal to reassembly, retiring twice, the zero / non-finite load
conventions and the hierarchy variants, each with the reference's tolerance.
"""
import numpy as np
import pytest

from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED
from tests import problems as P

pytestmark = pytest.mark.gpu


def shear(L):
    return np.hypot(L["V1"], L["V2"])


def moment(L):
    return np.hypot(L["M1"], L["M2"])


@pytest.mark.parametrize("precision", [SVXG_FP64, SVXG_MIXED], ids=["fp64", "mixed"])
def test_cantilever_carries_its_weight(precision, mats, ctx):
    # test_stress.cpp:59-87
    n, L, m = 12, 0.5, 300.0
    nodes, bonds = P.cantilever(n, L, m)
    p = stress.StressProblem(nodes, bonds, mats, ctx)
    assert p.assemble(stress.default_options(amg_min_nodes=4, precision=precision))
    u, r = p.solve(P.gravity(nodes), rtol=1e-10, maxit=500)
    assert r["converged"]
    loads = p.bond_loads(u)
    W = 9.81 * m * n
    M = sum(9.81 * m * (i + 0.5) * L for i in range(n))
    assert shear(loads[0]) == pytest.approx(W, rel=1e-6)
    assert moment(loads[0]) == pytest.approx(M, rel=1e-6)
    assert abs(loads[0]["N"]) < 1e-6 * W
    assert shear(loads[n - 1]) == pytest.approx(9.81 * m, rel=1e-6)
    assert moment(loads[n - 1]) == pytest.approx(9.81 * m * 0.5 * L, rel=1e-6)
    phi, _ = p.bond_checks(u)
    assert phi[0] > 50.0 * phi[n - 1]


def test_multigrid_and_block_jacobi_agree(mats, ctx):
    # test_stress.cpp:89-138
    nodes, bonds = P.cube_block(8, 0.4)
    A = stress.StressProblem(nodes, bonds, mats, ctx)
    B = stress.StressProblem(nodes, bonds, mats, ctx)
    assert A.assemble(stress.default_options(amg_min_nodes=16))
    assert B.assemble(stress.default_options(amg_min_nodes=1 << 30))
    ua, ra = A.solve(P.gravity(nodes), rtol=1e-9, maxit=2000)
    ub, rb = B.solve(P.gravity(nodes), rtol=1e-9, maxit=5000)
    assert ra["converged"] and rb["converged"]
    assert ra["iters"] < rb["iters"]
    assert np.abs(ua - ub).max() <= 1e-6 * np.abs(ua).max()
    loads = A.bond_loads(ua)
    R = loads["N"][bonds["b"] < 0].sum()
    assert -R == pytest.approx(9.81 * 400.0 * 512, rel=1e-5)


def test_free_chain_uniform_tension(mats, ctx):
    # test_stress.cpp:140-152
    nodes, bonds = P.cantilever(9, 0.5, 100.0)
    bonds = bonds[1:].copy()
    nodes["fixed"][4] = 1
    p = stress.StressProblem(nodes, bonds, mats, ctx)
    assert p.assemble(stress.default_options())
    f = np.zeros(6 * 9)
    f[0], f[6 * 8] = -1000.0, 1000.0
    u, r = p.solve(f, rtol=1e-10, maxit=500)
    assert r["converged"]
    assert np.allclose(p.bond_loads(u)["N"], 1000.0, rtol=1e-6)


def test_remove_bond_in_place_equals_reassembly(mats, ctx):
    # test_stress.cpp:154-174
    nodes, bonds = P.cantilever(10, 0.5, 200.0)
    bonds = P.add_bond(nodes, bonds, 6, -1, [0, 0, -1], nodes[6]["c"] - [0, 0, 0.25], 0.5)
    so = stress.default_options(amg_min_nodes=1 << 30)
    A = stress.StressProblem(nodes, bonds, mats, ctx)
    assert A.assemble(so)
    A.remove_bond(3)
    b2 = bonds.copy()
    b2["broken"][3] = 1
    B = stress.StressProblem(nodes, b2, mats, ctx)
    assert B.assemble(so)
    ua, _ = A.solve(P.gravity(nodes), rtol=1e-11, maxit=2000)
    ub, _ = B.solve(P.gravity(nodes), rtol=1e-11, maxit=2000)
    assert np.abs(ua - ub).max() <= 1e-6 * np.abs(ub).max()
    # (and the multigrid path keeps its stale hierarchy and still converges to the same answer)
    C = stress.StressProblem(nodes, bonds, mats, ctx)
    assert C.assemble(stress.default_options(amg_min_nodes=4))
    C.remove_bond(3)
    assert not C.info()["preconditioner_current"]
    uc, rc = C.solve(P.gravity(nodes), rtol=1e-11, maxit=2000)
    assert rc["converged"] and np.abs(uc - ub).max() <= 1e-6 * np.abs(ub).max()


def test_retiring_twice_never_subtracts_twice(mats, ctx):
    # test_stress.cpp:238-276
    nodes, bonds = P.cantilever(12, 0.5, 200.0)
    for i in (3, 6, 9):
        bonds = P.add_bond(nodes, bonds, i, -1, [0, 0, -1], nodes[i]["c"] - [0, 0, 0.25], 0.5)
    so = stress.default_options(amg_min_nodes=1 << 30)
    A = stress.StressProblem(nodes, bonds, mats, ctx)
    assert A.assemble(so)
    A.retire_nodes([10, 11])
    for k, b in enumerate(A.bonds):
        if b["a"] == 10 and b["b"] == 11:
            A.set_bond_broken(k, False)
    A.retire_nodes([1])
    n2 = nodes.copy()
    n2["gone"][[1, 10, 11]] = 1
    B = stress.StressProblem(n2, bonds, mats, ctx)
    assert B.assemble(so)
    f = P.gravity_live(n2)
    ua, ra = A.solve(f, rtol=1e-11, maxit=2000)
    ub, rb = B.solve(f, rtol=1e-11, maxit=2000)
    assert ra["converged"] and not ra["breakdown"] and rb["converged"]
    live = np.repeat(n2["gone"] == 0, 6)
    assert np.abs(ua - ub)[live].max() <= 1e-6 * np.abs(ub)[live].max()
    inner = [k for k, b in enumerate(A.bonds) if b["a"] == 10 and b["b"] == 11]
    assert inner and not A.bonds["broken"][inner[0]]
    assert A.assemble(so)
    assert not A.bonds["broken"][inner[0]]


def test_retired_rows_do_not_slow_the_multigrid(mats, ctx):
    # test_stress.cpp:278-291
    so = stress.default_options(amg_min_nodes=16)
    nw, bw = P.cube_block(14, 0.4, 9, True)
    no, bo = P.cube_block(14, 0.4, 9, False)
    W = stress.StressProblem(nw, bw, mats, ctx)
    O = stress.StressProblem(no, bo, mats, ctx)
    assert W.assemble(so) and O.assemble(so)
    _, rw = W.solve(P.gravity_live(nw), rtol=1e-8, maxit=500)
    _, ro = O.solve(P.gravity_live(no), rtol=1e-8, maxit=500)
    assert rw["converged"] and ro["converged"]
    assert rw["iters"] <= ro["iters"] + 4


@pytest.mark.parametrize("precision", [SVXG_FP64, SVXG_MIXED], ids=["fp64", "mixed"])
def test_zero_and_non_finite_loads(precision, mats, ctx):
    # test_stress.cpp:293-305
    nodes, bonds = P.cantilever(6, 0.5, 100.0)
    p = stress.StressProblem(nodes, bonds, mats, ctx)
    assert p.assemble(stress.default_options(precision=precision))
    u = np.ones(6 * 6)
    f = np.zeros(6 * 6)
    u, r0 = p.solve(f, u, rtol=1e-9, maxit=100, warm=True)
    assert r0["converged"] and np.all(u == 0.0)
    f[2] = np.nan
    u, r1 = p.solve(f, rtol=1e-9, maxit=100)
    assert not r1["converged"] and r1["breakdown"]


def test_hierarchy_variants_solve_alike(mats, ctx):
    # test_stress.cpp:307-381 (agreement within 1e-5 of max |u|)
    nodes, bonds = P.cube_block(8, 0.4)
    f = P.gravity(nodes)
    sols = []
    for kw in (dict(), dict(smoothed=0), dict(reaggregate=0), dict(coarsen_dense=0, reaggregate=0)):
        p = stress.StressProblem(nodes, bonds, mats, ctx)
        assert p.assemble(stress.default_options(**kw))
        u, r = p.solve(f, rtol=1e-8, maxit=500)
        assert r["converged"], kw
        sols.append(u)
    for u in sols[1:]:
        assert np.abs(u - sols[0]).max() <= 1e-5 * np.abs(sols[0]).max()


def test_begin_iterate_resume(mats, ctx):
    """A solve spread over calls (stress.cpp:362-475) reaches the one-call answer."""
    g = gen.tower()
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    assert p.assemble(stress.default_options())
    u1, r1 = p.solve(g.f, rtol=1e-8, maxit=500)
    p.begin(g.f)
    total = 0
    for _ in range(100):
        r = p.iterate(3, 1e-8)
        total += r["iters"]
        if r["converged"]:
            break
    assert r["converged"] and total == r1["iters"]
    u2 = p.current()
    assert np.array_equal(u1, u2)
