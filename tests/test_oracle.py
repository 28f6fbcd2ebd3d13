"""The oracle (oracle/oracle.cpp, a CPU restatement of the reference's solve) pinned three ways:

1. bit for bit against the reference's own sources compiled by oracle/Makefile (oracle/_ref), where
   those build (this container);
2. against the committed golden fixtures tests/golden/*.npz that tests/golden/make_golden.py wrote
   from oracle/_ref (everywhere, the GPU box included);
3. against the closed forms and known answers of tests/core/test_stress.cpp (cantilever statics,
   chain tension, support reactions, failure modes) and a brute-force dense solve.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from oracle import CpuProblem, reference_options
from opus_destruct_2_amd import gen
from tests import problems as P

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def golden_problem(z):
    """(nodes, bonds, materials, options) of a fixture (structs stored as raw bytes)."""
    from opus_destruct_2_amd.abi import BOND, MATERIAL, NODE
    nodes = np.frombuffer(z["nodes"].tobytes(), NODE).copy()
    bonds = np.frombuffer(z["bonds"].tobytes(), BOND).copy()
    fmats = np.frombuffer(z["mats"].tobytes(), MATERIAL).copy()
    opt = reference_options(**{str(k): (int(v) if k not in ("strength", "coarse_scale", "rtol") else float(v))
                               for k, v in zip(z["opt_keys"], z["opt_vals"])})
    return nodes, bonds, fmats, opt


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


# ---- closed forms (test_stress.cpp:59-87, 134-152, 176-198) -------------------------------------------

def test_cantilever_statics(mats):
    n, L, m = 12, 0.5, 300.0
    nodes, bonds = P.cantilever(n, L, m)
    p = CpuProblem(nodes, bonds, mats)
    assert p.assemble(reference_options(amg_min_nodes=4))
    u, r = p.solve(P.gravity(nodes), rtol=1e-10, maxit=500)
    assert r["converged"]
    root = p.bond_load(0, u)
    W = 9.81 * m * n
    M = sum(9.81 * m * (i + 0.5) * L for i in range(n))
    assert np.hypot(root["V1"], root["V2"]) == pytest.approx(W, rel=1e-6)
    assert np.hypot(root["M1"], root["M2"]) == pytest.approx(M, rel=1e-6)
    assert abs(root["N"]) < 1e-6 * W
    tip = p.bond_load(n - 1, u)
    assert np.hypot(tip["V1"], tip["V2"]) == pytest.approx(9.81 * m, rel=1e-6)
    assert np.hypot(tip["M1"], tip["M2"]) == pytest.approx(9.81 * m * 0.5 * L, rel=1e-6)


def test_chain_tension(mats):
    nodes, bonds = P.cantilever(9, 0.5, 100.0)
    bonds = bonds[1:].copy()
    nodes["fixed"][4] = 1
    p = CpuProblem(nodes, bonds, mats)
    assert p.assemble(reference_options())
    f = np.zeros(6 * 9)
    f[0] = -1000.0
    f[6 * 8] = 1000.0
    u, r = p.solve(f, rtol=1e-10, maxit=500)
    assert r["converged"]
    for b in range(len(bonds)):
        assert p.bond_load(b, u)["N"] == pytest.approx(1000.0, rel=1e-6)


def test_support_reactions_and_multigrid_vs_jacobi(mats):
    nodes, bonds = P.cube_block(8, 0.4)
    a = CpuProblem(nodes, bonds, mats)
    b = CpuProblem(nodes, bonds, mats)
    assert a.assemble(reference_options(amg_min_nodes=16))
    assert b.assemble(reference_options(amg_min_nodes=1 << 30))
    ua, ra = a.solve(P.gravity(nodes), rtol=1e-9, maxit=2000)
    ub, rb = b.solve(P.gravity(nodes), rtol=1e-9, maxit=5000)
    assert ra["converged"] and rb["converged"] and ra["iters"] < rb["iters"]
    assert np.abs(ua - ub).max() <= 1e-6 * np.abs(ua).max()
    R = sum(a.bond_load(k, ua)["N"] for k in range(len(bonds)) if bonds[k]["b"] < 0)
    assert -R == pytest.approx(9.81 * 400.0 * 512, rel=1e-5)


def test_failure_modes_known_answers(mats):
    nodes, bonds = P.cantilever(2, 0.5, 1.0)
    b = bonds[1]
    S = oracle.bond_strength(b, 1.0, mats)
    area = b["area"]
    ft, fb, fc, coh, mu = S
    assert oracle.bond_utilization(b, {"N": 1.01 * ft * area}, 1.0, mats) == (pytest.approx(1.01), 1)
    assert oracle.bond_utilization(b, {"N": -1.2 * fc * area}, 1.0, mats) == (pytest.approx(1.2), 2)
    assert oracle.bond_utilization(b, {"V1": coh * area}, 1.0, mats) == (pytest.approx(1.5), 3)
    L = {"V1": coh * area, "N": -area * coh / mu}
    assert oracle.bond_utilization(b, L, 1.0, mats)[0] == pytest.approx(0.75)
    assert oracle.bond_utilization(b, L, 2.0, mats)[0] == pytest.approx(1.0)


def test_solution_matches_dense_direct_solve(mats):
    """u of a small irregular piece equals numpy's dense solve of the same K (built from the blocks)."""
    g = gen.piece(60, pinned=False, seed=3)
    p = CpuProblem(g.nodes, g.bonds, mats)
    assert p.assemble(reference_options(amg_min_nodes=0))
    rowptr, col, val = p.matrix()
    n = len(rowptr) - 1
    K = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for k in range(rowptr[i], rowptr[i + 1]):
            j = col[k]
            K[6 * i:6 * i + 6, 6 * j:6 * j + 6] = val[k]
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    assert np.linalg.eigvalsh(0.5 * (K + K.T)).min() > 0
    u, r = p.solve(g.f, rtol=1e-12, maxit=1000)
    ud = np.linalg.solve(K, g.f)
    assert np.abs(u - ud).max() <= 1e-8 * np.abs(ud).max()


def test_rigid_motion_loads_nothing(mats):
    """bond_load of a rigid motion of a free chain is zero (K's null space)."""
    nodes, bonds = P.cantilever(5, 0.5, 100.0)
    bonds = bonds[1:].copy()
    p = CpuProblem(nodes, bonds, mats)
    p.assemble(reference_options())
    U, Th = np.array([1e-3, -2e-3, 5e-4]), np.array([1e-4, 3e-4, -2e-4])
    u = np.zeros(6 * 5)
    for i in range(5):
        u[6 * i:6 * i + 3] = U + np.cross(Th, nodes[i]["c"])
        u[6 * i + 3:6 * i + 6] = Th
    for b in range(len(bonds)):
        L = p.bond_load(b, u)
        assert max(abs(L[k]) for k in ("N", "V1", "V2", "T", "M1", "M2")) < 1e-3


# ---- the 6x6 block inverse (amg.cpp:105-135) and a level's diagonal inverse (:227-244) ---------------

def _blocks(seed):
    """SPD stiffness-like blocks, general blocks whose elimination pivots, and badly scaled ones."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(40):
        m = rng.standard_normal((6, 6))
        out.append(m @ m.T + 0.1 * np.eye(6))
        out.append(rng.standard_normal((6, 6)))
        out.append(np.diag(10.0 ** rng.uniform(-6, 9, 6)) @ rng.standard_normal((6, 6)))
    return out


def test_inv6_closed_forms():
    # diagonal: exact reciprocals; a permutation: its transpose, exactly; a rigid-motion map
    # [[I, -[r]x], [0, I]]: [[I, [r]x], [0, I]] exactly (its entries are small integers)
    d = np.array([2.0, 4.0, 0.5, 8.0, 1.0, 0.25])
    ok, inv = oracle.inv6(np.diag(d))
    assert ok and np.array_equal(inv, np.diag(1.0 / d))
    perm = np.eye(6)[[3, 0, 5, 1, 4, 2]]
    ok, inv = oracle.inv6(perm)
    assert ok and np.array_equal(inv, perm.T)
    r = np.array([1.0, -2.0, 3.0])
    X = np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
    Pm = np.eye(6)
    Pm[:3, 3:] = -X
    ok, inv = oracle.inv6(Pm)
    want = np.eye(6)
    want[:3, 3:] = X
    assert ok and np.array_equal(inv, want)
    # A A^-1 = I on well-conditioned blocks; against numpy's LU solve
    for a in _blocks(3)[:60]:
        ok, inv = oracle.inv6(a)
        assert ok
        assert np.abs(a @ inv - np.eye(6)).max() <= 1e-9 * np.linalg.cond(a)
        assert np.allclose(inv, np.linalg.inv(a), rtol=1e-8 * np.linalg.cond(a), atol=0)


def test_inv6_singular_rules():
    # zero block; an exactly rank-5 block; a pivot at 1e-16 of the largest entry: all singular.
    # At 1e-14 (above the 1e-15 pivot tolerance) it inverts.
    assert oracle.inv6(np.zeros((6, 6)))[0] is False
    a = np.eye(6)
    a[5, 5] = 0.0
    assert oracle.inv6(a)[0] is False
    a[5, 5] = 1e-16
    assert oracle.inv6(a)[0] is False
    a[5, 5] = 1e-14
    ok, inv = oracle.inv6(a)
    assert ok and inv[5, 5] == 1e14
    # the level's regularisation: a singular block gains 1e-9 x sum |diag| on its diagonal
    a = np.eye(6)
    a[5, 5] = 0.0
    got = oracle.diag_inverse(a)
    assert got is not None
    assert got[5, 5] == 1.0 / 5e-9 and got[0, 0] == 1.0 / (1.0 + 5e-9)
    assert oracle.diag_inverse(np.zeros((6, 6)))[0, 0] == 1.0  # (trace 0: +1 on the diagonal)


# ---- the compiled reference (where it builds) ------------------------------------------------------

ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def _cases():
    out = []
    nodes, bonds = P.cantilever(12, 0.5, 300.0)
    out.append(("cantilever", nodes, bonds, P.gravity(nodes), reference_options(amg_min_nodes=4), 1e-10, 500))
    nodes, bonds = P.cube_block(8, 0.4)
    out.append(("block8_amg", nodes, bonds, P.gravity(nodes), reference_options(amg_min_nodes=16), 1e-9, 2000))
    out.append(("block8_jacobi", nodes, bonds, P.gravity(nodes), reference_options(amg_min_nodes=1 << 30), 1e-9, 5000))
    nodes, bonds = P.cube_block(16, 0.4)
    out.append(("block16", nodes, bonds, P.gravity(nodes), reference_options(), 1e-8, 500))
    out.append(("block16_unsmoothed", nodes, bonds, P.gravity(nodes), reference_options(smoothed=0), 1e-8, 500))
    out.append(("block16_dense", nodes, bonds, P.gravity(nodes), reference_options(reaggregate=0), 1e-8, 500))
    nodes, bonds = P.cube_block(14, 0.4, 9, True)
    out.append(("retired", nodes, bonds, P.gravity_live(nodes), reference_options(amg_min_nodes=16), 1e-8, 500))
    g = gen.tower()
    out.append(("tower", g.nodes, g.bonds, g.f, reference_options(), 1e-8, 5000))
    for k in range(3):
        g = gen.piece(200 + 150 * k, pinned=bool(k & 1), seed=k + 11)
        out.append((f"piece{k}", g.nodes, g.bonds, g.f, reference_options(amg_min_nodes=0), 1e-8, 5000))
    return out


@ref
@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_oracle_bit_identical_to_reference(case, mats):
    name, nodes, bonds, f, opt, rtol, maxit = case
    a = CpuProblem(nodes, bonds, mats, "oracle")
    b = CpuProblem(nodes, bonds, mats, "ref")
    assert a.assemble(opt) and b.assemble(opt)
    ua, ra = a.solve(f, rtol=rtol, maxit=maxit)
    ub, rb = b.solve(f, rtol=rtol, maxit=maxit)
    assert ra == rb
    assert np.array_equal(_bits(ua), _bits(ub))
    ia, ib = a.info(), b.info()
    for k in ("nfree", "blocks", "work_per_iteration", "assembly_work", "block_jacobi"):
        assert ia[k] == ib[k], k
    pa, ma = a.bond_checks(ua)
    pb, mb = b.bond_checks(ub)
    assert np.array_equal(_bits(pa), _bits(pb)) and np.array_equal(ma, mb)


@ref
def test_oracle_incremental_updates_bit_identical(mats):
    """remove_bond in place, retire_nodes twice (test_stress.cpp:154-174, 238-276) on both."""
    nodes, bonds = P.cantilever(12, 0.5, 200.0)
    for i in (3, 6, 9):
        bonds = P.add_bond(nodes, bonds, i, -1, [0, 0, -1], nodes[i]["c"] - [0, 0, 0.25], 0.5)
    gone = np.zeros(len(nodes), bool)
    gone[[1, 10, 11]] = True
    f = P.gravity(nodes)
    f[2::6][gone] = 0.0
    res = {}
    for kind in ("oracle", "ref"):
        p = CpuProblem(nodes, bonds, mats, kind)
        assert p.assemble(reference_options(amg_min_nodes=1 << 30))
        p.retire_nodes([10, 11])
        for k, b in enumerate(p.get_bonds()):
            if b["a"] == 10 and b["b"] == 11:
                p.set_bond_broken(k, False)
        p.retire_nodes([1])
        p.remove_bond(5)
        u, r = p.solve(f, rtol=1e-11, maxit=2000)
        res[kind] = (u, r, p.get_bonds())
    assert res["oracle"][1] == res["ref"][1]
    assert np.array_equal(_bits(res["oracle"][0]), _bits(res["ref"][0]))
    assert np.array_equal(res["oracle"][2]["broken"], res["ref"][2]["broken"])
    assert np.array_equal(res["oracle"][2]["in_k"], res["ref"][2]["in_k"])


# ---- committed fixtures (written from oracle/_ref by tests/golden/make_golden.py) --------------------

@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))), ids=os.path.basename)
def test_oracle_matches_golden(path, mats):
    z = np.load(path, allow_pickle=False)
    nodes, bonds, fmats, opt = golden_problem(z)
    p = CpuProblem(nodes, bonds, fmats)
    assert p.assemble(opt)
    u, r = p.solve(z["f"], rtol=float(z["rtol"]), maxit=int(z["maxit"]))
    assert r["iters"] == int(z["iters"]) and r["converged"] == bool(z["converged"])
    assert np.array_equal(_bits(u), _bits(z["u"]))
    phi, mode = p.bond_checks(u)
    assert np.array_equal(_bits(phi), _bits(z["phi"])) and np.array_equal(mode, z["mode"])


def test_golden_fixtures_present():
    assert len(glob.glob(os.path.join(GOLDEN, "*.npz"))) >= 4


def test_extend_warm_start_parity_and_rigid_motion():
    """extend_warm_start: the library's, the restatement's and the reference's agree bit for bit;
    a rigid motion of the older nodes is carried exactly to the appended ones they reach."""
    from opus_destruct_2_amd import gen, stress
    import oracle
    g = gen.piece(300, pinned=False, seed=8)
    c = g.nodes["c"]
    t, w = np.array([1e-3, -2e-3, 5e-4]), np.array([2e-4, -1e-4, 3e-4])
    u = np.zeros((g.n, 6))
    u[:, :3] = t + np.cross(w, c)
    u[:, 3:] = w
    u = u.reshape(-1)
    first = 180
    u0 = u.copy()
    u0[6 * first:] = 0.0
    bonds = g.bonds.copy()
    bonds["broken"][::7] = 1  # (broken bonds carry nothing)
    got = {"product": stress.extend_warm_start(g.nodes, bonds, u0, first),
           "oracle": oracle.extend_warm_start(g.nodes, bonds, u0, first)}
    if oracle.ref_available():
        got["reference"] = oracle.extend_warm_start(g.nodes, bonds, u0, first, kind="ref")
    for k, v in got.items():
        assert np.array_equal(v.view(np.uint64), got["product"].view(np.uint64)), k
    ue = got["product"].reshape(-1, 6)
    reached = np.abs(ue[first:]).sum(axis=1) > 0
    assert reached.sum() > 0.5 * (g.n - first)
    ur = u.reshape(-1, 6)[first:]
    assert np.allclose(ue[first:][reached], ur[reached], rtol=1e-12, atol=1e-15)
    assert np.array_equal(ue[:first], u.reshape(-1, 6)[:first])


@ref
def test_inv6_bit_identical_to_reference():
    """The oracle's inv6 and the level regularisation against blk6::inv6 itself, bit for bit (and its
    singular verdicts), on well-conditioned, pivoting, badly scaled and singular blocks."""
    blocks = _blocks(5)
    a = np.eye(6)
    a[2, 2] = 0.0
    blocks += [np.zeros((6, 6)), a, np.ones((6, 6))]
    for k, b in enumerate(blocks):
        ok_o, inv_o = oracle.inv6(b, "oracle")
        ok_r, inv_r = oracle.inv6(b, "ref")
        assert ok_o == ok_r, k
        if ok_o:
            assert np.array_equal(_bits(inv_o), _bits(inv_r)), k
        do, dr = oracle.diag_inverse(b, "oracle"), oracle.diag_inverse(b, "ref")
        assert (do is None) == (dr is None), k
        if do is not None:
            assert np.array_equal(_bits(do), _bits(dr)), k
