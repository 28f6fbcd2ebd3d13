"""Parity of the device library with the oracle (the reference's algorithm on the CPU).

Bars (the path's contract):
  K (assembly) and the FP64 block product: bit-identical to the reference's (same operation order).
  bond loads and utilisations from a given u: bit-identical.
  displacements u: max |u - u_ref| <= 1e-6 max |u_ref| (FP64 path), 1e-5 (mixed path), both solves
  taken to a relative residual of 1e-10 so that the comparison measures the solvers, not the
  stopping points;
  utilisations phi: |phi - phi_ref| <= 1e-6 max(1, |phi_ref|) (FP64), 1e-5 (mixed);
  failed bonds {phi >= 1}: identical except bonds within 1e-4 of phi = 1.
"""
import numpy as np
import pytest

import oracle
from oracle import CpuProblem, reference_options
from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED
from tests import problems as P

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def cases():
    out = []
    nodes, bonds = P.cantilever(12, 0.5, 300.0)
    out.append(("cantilever12", nodes, bonds, P.gravity(nodes), dict(amg_min_nodes=4)))
    nodes, bonds = P.cantilever(12, 0.5, 300.0)
    out.append(("cantilever12_jacobi", nodes, bonds, P.gravity(nodes), dict()))
    nodes, bonds = P.cube_block(8, 0.4)
    out.append(("block8", nodes, bonds, P.gravity(nodes), dict(amg_min_nodes=16)))
    nodes, bonds = P.cube_block(10, 0.4, 7, True)
    out.append(("block10_retired", nodes, bonds, P.gravity_live(nodes), dict(amg_min_nodes=16)))
    g = gen.tower()
    out.append(("tower8x8x20", g.nodes, g.bonds, g.f, dict()))
    g = gen.piece(400, pinned=True, seed=7)
    out.append(("piece400_pinned", g.nodes, g.bonds, g.f, dict(amg_min_nodes=0)))
    g = gen.piece(700, pinned=False, seed=9)
    out.append(("piece700", g.nodes, g.bonds, g.f, dict(amg_min_nodes=0)))
    # (the single-large-structure generator at a small size: 15 % voids, materials mixed per column)
    g = gen.block(12, 12, 10, 0.5, 0.15, seed=5)
    out.append(("block12x12x10_voids", g.nodes, g.bonds, g.f, dict()))
    return out


CASES = cases()
IDS = [c[0] for c in CASES]
# the solve's precision: plain FP64; mixed (FP64 K products, FP32 V-cycle); mixed with FP32 K
# products too, inside FP64 iterative refinement; mixed with the FP16 V-cycle copies, and with them
# and the lean hierarchy (the settings of graphs of 16384+ free nodes, forced on these); FP64 lean
PRECISIONS = [(SVXG_FP64, {}), (SVXG_MIXED, {}), (SVXG_MIXED, {"mixed_fp64_rows": 0}),
              (SVXG_MIXED, {"vcycle_fp16_rows": 0}), (SVXG_MIXED, {"vcycle_fp16_rows": 0, "lean_rows": 1}),
              (SVXG_FP64, {"lean_rows": 1})]
PREC_IDS = ["fp64", "mixed", "mixed_ir", "mixed_vc16", "mixed_vc16_lean", "fp64_lean"]


def failed_sets_agree(phi, phi_ref, band=1e-4):
    a, b = phi >= 1.0, phi_ref >= 1.0
    diff = a != b
    near = (np.abs(phi - 1.0) <= band) | (np.abs(phi_ref - 1.0) <= band)
    return not np.any(diff & ~near)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_matrix_bit_identical(case, mats, ctx):
    name, nodes, bonds, f, kw = case
    d = stress.StressProblem(nodes, bonds, mats, ctx)
    assert d.assemble(stress.default_options(**kw))
    o = CpuProblem(nodes, bonds, mats)
    assert o.assemble(reference_options(**kw))
    rd, cd, vd = d.matrix()
    ro, co, vo = o.matrix()
    assert np.array_equal(rd, ro) and np.array_equal(cd, co)
    assert np.array_equal(bits(vd), bits(vo))
    assert np.array_equal(d.bonds["in_k"], o.get_bonds()["in_k"])


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("prec", [SVXG_FP64, SVXG_MIXED], ids=["fp64", "mixed"])
def test_diag_inverse_bit_identical(case, prec, mats, ctx):
    """Level 0's inverse diagonal blocks (Amg::finalize, amg.cpp:227-244; block-Jacobi's, stress.cpp:
    149-157) against the reference's inv6 applied to the device's own K: bit for bit where the level
    holds FP64 blocks (the FP64 path), the FP32 rounding of those bits where it holds FP32 copies."""
    name, nodes, bonds, f, kw = case
    d = stress.StressProblem(nodes, bonds, mats, ctx)
    assert d.assemble(stress.default_options(precision=prec, **kw))
    rowptr, col, val = d.matrix()
    dinv, held64 = d.diag_inverse(0)
    assert held64 == (prec == SVXG_FP64 or name.endswith("_jacobi"))
    kind = "ref" if oracle.ref_available() else "oracle"
    want = np.zeros_like(dinv)
    for i in range(len(rowptr) - 1):
        assert col[rowptr[i]] == i
        w = oracle.diag_inverse(val[rowptr[i]], kind)
        assert w is not None, (name, i)
        want[i] = w
    if held64:
        assert np.array_equal(bits(dinv), bits(want)), name
    else:
        assert np.array_equal(bits(dinv), bits(want.astype(np.float32).astype(np.float64))), name


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_block_product_bit_identical(case, mats, ctx):
    name, nodes, bonds, f, kw = case
    d = stress.StressProblem(nodes, bonds, mats, ctx)
    assert d.assemble(stress.default_options(**kw))
    o = CpuProblem(nodes, bonds, mats)
    o.assemble(reference_options(**kw))
    n6 = 6 * d.info()["nfree"]
    x = gen.uniform(17, n6) - 0.5
    y = d.apply_matrix(x)
    assert np.array_equal(bits(y), bits(o.apply_matrix(x)))
    y32 = d.apply_matrix(x, fp32=True)
    assert np.abs(y32 - y).max() <= 1e-5 * np.abs(y).max()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("prec", PRECISIONS, ids=PREC_IDS)
def test_solution_matches_oracle(case, prec, mats, ctx):
    name, nodes, bonds, f, kw = case
    precision, pkw = prec
    tol = 1e-6 if precision == SVXG_FP64 else 1e-5
    o = CpuProblem(nodes, bonds, mats)
    assert o.assemble(reference_options(**kw))
    uo, ro = o.solve(f, rtol=1e-10, maxit=20000)
    assert ro["converged"]
    d = stress.StressProblem(nodes, bonds, mats, ctx)
    assert d.assemble(stress.default_options(precision=precision, **kw, **pkw))
    ud, rd = d.solve(f, rtol=1e-10, maxit=20000)
    assert rd["converged"], rd
    assert rd["rel_res"] <= 1e-10
    scale = np.abs(uo).max()
    assert np.abs(ud - uo).max() <= tol * scale, (np.abs(ud - uo).max() / scale)
    phi_o, _ = o.bond_checks(uo)
    phi_d, _ = d.bond_checks(ud)
    assert np.all(np.abs(phi_d - phi_o) <= tol * np.maximum(1.0, np.abs(phi_o)))
    assert failed_sets_agree(phi_d, phi_o)


@pytest.mark.parametrize("case", CASES[2:5], ids=IDS[2:5])
def test_bond_checks_bit_identical(case, mats, ctx):
    name, nodes, bonds, f, kw = case
    o = CpuProblem(nodes, bonds, mats)
    o.assemble(reference_options(**kw))
    uo, _ = o.solve(f, rtol=1e-8, maxit=5000)
    d = stress.StressProblem(nodes, bonds, mats, ctx)
    loads = d.bond_loads(uo)
    lo = o.bond_loads(uo)
    for k in ("N", "V1", "V2", "T", "M1", "M2"):
        assert np.array_equal(bits(loads[k]), bits(lo[k])), k
    for frag in (1.0, 2.5):
        phi_d, mode_d = d.bond_checks(uo, frag)
        phi_o, mode_o = o.bond_checks(uo, frag)
        assert np.array_equal(bits(phi_d), bits(phi_o)) and np.array_equal(mode_d, mode_o)


def test_resident_solution_checks(mats, ctx):
    """After a solve the checks run on the device-resident u (no upload): the same answer."""
    g = gen.tower()
    d = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    d.assemble(stress.default_options())
    u, r = d.solve(g.f, rtol=1e-8)
    assert r["converged"]
    a = d.bond_checks(None)
    b = d.bond_checks(u)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_vcycle_fp16_copies(mats, ctx):
    """The V-cycle's FP16 block copies (the mixed default for large graphs) against its FP32 copies: the same
    iterations within a few, the same answer within the mixed bar, a symmetric preconditioner; a
    change of the option needs a new setup."""
    g = gen.block(16, 16, 12, 0.5, 0.15, seed=3)
    us, its = [], []
    for fp16 in (1, 0):
        opts = stress.default_options(precision=SVXG_MIXED, vcycle_fp16_rows=0 if fp16 else 1 << 30)
        p = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts)
        assert p.assemble(opts)
        u, r = p.solve(g.f, rtol=1e-10, maxit=2000)
        assert r["converged"]
        us.append(u)
        its.append(r["iters"])
        n6 = 6 * p.info()["nfree"]
        rng = np.random.default_rng(5)
        r1, r2 = rng.standard_normal(n6), rng.standard_normal(n6)
        z1, z2 = p.apply_preconditioner(r1), p.apply_preconditioner(r2)
        assert abs(r2 @ z1 - r1 @ z2) <= 1e-5 * np.linalg.norm(r2) * np.linalg.norm(z1)
        if fp16:
            p.set_options(stress.default_options(precision=SVXG_MIXED, vcycle_fp16_rows=1 << 30))
            assert not p.info()["assembled"]
        p.close()
    assert abs(its[0] - its[1]) <= 3
    scale = np.abs(us[1]).max()
    assert np.abs(us[0] - us[1]).max() <= 1e-6 * scale


@pytest.mark.parametrize("case", [CASES[3], CASES[7]], ids=[IDS[3], IDS[7]])
def test_matrix_after_removals_bit_identical(case, mats, ctx):
    """K after bond removals in place (StressProblem::remove_bond, stress.cpp:260-284): many bonds,
    neighbours among them (their diagonal blocks take several products, in the list's order), in
    two calls and in one batch of two problems, bit for bit the oracle's removing them one by one."""
    name, nodes, bonds, f, kw = case
    rng = np.random.default_rng(11)
    live = np.flatnonzero(bonds["broken"] == 0)
    cut1 = rng.choice(live, size=len(live) // 8, replace=False)
    cut2 = rng.choice(np.setdiff1d(live, cut1), size=len(live) // 10, replace=False)
    o = CpuProblem(nodes, bonds, mats)
    assert o.assemble(reference_options(**kw))
    for b in list(cut1) + list(cut2):
        o.remove_bond(int(b))
    ro, co, vo = o.matrix()
    d = stress.StressProblem(nodes, bonds, mats, ctx)
    assert d.assemble(stress.default_options(**kw))
    d.remove_bonds(cut1)
    d.remove_bonds(cut2)
    rd, cd, vd = d.matrix()
    assert np.array_equal(rd, ro) and np.array_equal(cd, co)
    assert np.array_equal(bits(vd), bits(vo))
