"""Batches of uneven graphs (the reference solves its pieces one by one, world_pieces.cpp:1452-1463):
each graph's answer in a batch is bitwise its answer alone, and matches the oracle."""
import numpy as np
import pytest

from oracle import CpuProblem, reference_options
from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED

pytestmark = pytest.mark.gpu

SIZES = [16, 31, 48, 60, 100, 159, 161, 250, 400, 900, 2000, 3500]
# plain FP64; mixed (FP64 K products, FP32 V-cycle); mixed with FP32 K products in iterative refinement
PRECISIONS = [(SVXG_FP64, {}), (SVXG_MIXED, {}), (SVXG_MIXED, {"mixed_fp64_rows": 0})]
PREC_IDS = ["fp64", "mixed", "mixed_ir"]


def graphs():
    return [gen.piece(s, pinned=bool(k & 1), seed=100 + k) for k, s in enumerate(SIZES)]


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


@pytest.mark.parametrize("min_nodes", [0, 160], ids=["all_multigrid", "small_jacobi"])
@pytest.mark.parametrize("prec", PRECISIONS, ids=PREC_IDS)
def test_batch_equals_alone_bitwise(prec, min_nodes, mats, ctx):
    gs = graphs()
    opts = stress.default_options(precision=prec[0], amg_min_nodes=min_nodes, **prec[1])
    alone = []
    for g in gs:
        p = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts)
        assert p.assemble()
        u, r = p.solve(g.f, rtol=1e-9, maxit=5000)
        assert r["converged"], r
        alone.append((u, r, p.info()))
        p.close()
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    us, rs = stress.solve_batch(probs, [g.f for g in gs], rtol=1e-9, maxit=5000)
    for k, g in enumerate(gs):
        assert rs[k] == alone[k][1], (SIZES[k], rs[k], alone[k][1])
        assert probs[k].info()["level_rows"] == alone[k][2]["level_rows"]
        assert np.array_equal(bits(us[k]), bits(alone[k][0])), SIZES[k]


@pytest.mark.parametrize("prec", PRECISIONS, ids=PREC_IDS)
def test_batch_matches_oracle(prec, mats, ctx):
    gs = graphs()
    precision = prec[0]
    tol = 1e-6 if precision == SVXG_FP64 else 1e-5
    opts = stress.default_options(precision=precision, amg_min_nodes=0, **prec[1])
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    assert all(stress.assemble_batch(probs))
    us, rs = stress.solve_batch(probs, [g.f for g in gs], rtol=1e-10, maxit=5000)
    for k, g in enumerate(gs):
        o = CpuProblem(g.nodes, g.bonds, mats)
        assert o.assemble(reference_options(amg_min_nodes=0))
        uo, ro = o.solve(g.f, rtol=1e-10, maxit=20000)
        assert rs[k]["converged"] and ro["converged"]
        scale = np.abs(uo).max()
        assert np.abs(us[k] - uo).max() <= tol * scale, (SIZES[k], np.abs(us[k] - uo).max() / scale)
        phi_o, _ = o.bond_checks(uo)
        phi_d, mode_d = probs[k].bond_checks(None)
        assert np.all(np.abs(phi_d - phi_o) <= tol * np.maximum(1.0, np.abs(phi_o)))
        # (the caller's arrays, written over: the same bits)
        out = (np.full(len(phi_d), np.nan), np.full(len(phi_d), 7, np.uint8))
        phi_w, mode_w = probs[k].bond_checks(None, out=out)
        assert phi_w is out[0] and np.array_equal(phi_w.view(np.int64), phi_d.view(np.int64))
        assert np.array_equal(mode_w, mode_d)
        with pytest.raises(ValueError):
            probs[k].bond_checks(None, out=(np.zeros(len(phi_d), np.float32), mode_w))


def test_batch_warm_start_and_resident_loads(mats, ctx):
    """A warm start at the answer converges at once; a null load reuses the resident one."""
    gs = graphs()[:6]
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, ctx) for g in gs]
    us, rs = stress.solve_batch(probs, [g.f for g in gs], rtol=1e-9)
    us2, rs2 = stress.solve_batch(probs, [g.f for g in gs], us=us, rtol=1e-9, warm=True)
    for k in range(len(gs)):
        assert rs2[k]["converged"] and rs2[k]["iters"] <= 1


@pytest.mark.parametrize("prec", PRECISIONS, ids=PREC_IDS)
def test_resident_steps_equal_tiled_steps_bitwise(prec, mats):
    """Graphs stepped in one workgroup each give the bits the tiled launches give (the same device
    functions on the same tiles in the same order)."""
    gs = graphs()[:10]
    out = []
    for rows in (0, 4096):
        c = stress.Context(0)
        c.set_resident_rows(rows)
        opts = stress.default_options(precision=prec[0], amg_min_nodes=0 if rows else 160, **prec[1])
        opts.amg_min_nodes = 0
        probs = [stress.StressProblem(g.nodes, g.bonds, mats, c, options=opts) for g in gs]
        us, rs = stress.solve_batch(probs, [g.f for g in gs], rtol=1e-9, maxit=5000)
        us2, rs2 = stress.solve_batch(probs, [g.f for g in gs], us=us, rtol=1e-11, maxit=5000, warm=True)
        out.append((us, rs, us2, rs2))
        for p in probs:
            p.close()
        c.close()
    (a, ra, a2, ra2), (b, rb, b2, rb2) = out
    for k in range(len(gs)):
        assert ra[k] == rb[k] and ra2[k] == rb2[k], (SIZES[k], ra[k], rb[k])
        assert np.array_equal(bits(a[k]), bits(b[k])) and np.array_equal(bits(a2[k]), bits(b2[k])), SIZES[k]


@pytest.mark.parametrize("prec", PRECISIONS, ids=PREC_IDS)
def test_step_graphs_equal_kernel_launches_bitwise(prec, mats):
    """Solve steps replayed from a recorded HIP graph give the bits of the same kernels launched one
    by one: batched, and a single problem solved again and again (its batch, and so its graph,
    kept between calls, with other device work in between)."""
    gs = graphs()
    out = []
    for mode in (2, 0):  # (recorded from the first step, whatever the batch's size; never)
        c = stress.Context(0)
        c.set_step_graphs(mode)
        opts = stress.default_options(precision=prec[0], amg_min_nodes=0, **prec[1])
        probs = [stress.StressProblem(g.nodes, g.bonds, mats, c, options=opts) for g in gs]
        us, rs = stress.solve_batch(probs, [g.f for g in gs], rtol=1e-9, maxit=5000)
        p = probs[-1]
        seq = []
        for k in range(3):
            f = gs[-1].f * (1.0 + 0.5 * k)
            seq.append(p.solve(f, rtol=1e-9, maxit=5000))
            p.apply_matrix(gen.uniform(k, 6 * p.info()["nfree"]))
        out.append((us, rs, seq))
        for q in probs:
            q.close()
        c.close()
    (a, ra, sa), (b, rb, sb) = out
    for k in range(len(gs)):
        assert ra[k] == rb[k], (SIZES[k], ra[k], rb[k])
        assert np.array_equal(bits(a[k]), bits(b[k])), SIZES[k]
    for (ua, qa), (ub, qb) in zip(sa, sb):
        assert qa == qb and np.array_equal(bits(ua), bits(ub))


@pytest.mark.parametrize("prec", PRECISIONS[:2], ids=PREC_IDS[:2])
@pytest.mark.parametrize("step_graphs", [2, 1], ids=["recorded", "default"])
def test_batch_tail_compaction_bitwise(prec, step_graphs, mats):
    """A batch whose running graphs hold at most half of its level-0 tiles runs its launches over
    their tiles only (the tables cut at looks): 48 pieces of 86k nodes (2,200 tiles), the
    larger ones warm-started at their solutions (converged at the start), the rest cold; each
    piece's result and iterations are bitwise its own solve alone from the same start, with steps
    recorded as graphs (dropped when the tables are cut) and launched one by one."""
    sizes = [4000, 3000, 2500, 2000, 1500, 900, 400, 100] * 6
    gs = [gen.piece(s, pinned=bool(k & 1), seed=500 + k) for k, s in enumerate(sizes)]
    c = stress.Context(0)
    c.set_step_graphs(step_graphs)
    opts = stress.default_options(precision=prec[0], amg_min_nodes=0, **prec[1])
    u0 = []
    for k, g in enumerate(gs):
        if sizes[k] >= 2500:
            p = stress.StressProblem(g.nodes, g.bonds, mats, c, options=opts)
            assert p.assemble()
            u, r = p.solve(g.f, rtol=1e-9, maxit=5000)
            assert r["converged"], r
            u0.append(u)
            p.close()
        else:
            u0.append(np.zeros(6 * len(g.nodes)))
    alone = []
    for k, g in enumerate(gs):
        p = stress.StressProblem(g.nodes, g.bonds, mats, c, options=opts)
        us, rs = stress.solve_batch([p], [g.f], us=[u0[k]], rtol=1e-9, maxit=5000, warm=True)
        assert rs[0]["converged"], rs[0]
        alone.append((us[0], rs[0]))
        p.close()
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, c, options=opts) for g in gs]
    us, rs = stress.solve_batch(probs, [g.f for g in gs], us=u0, rtol=1e-9, maxit=5000, warm=True)
    for k in range(len(gs)):
        assert rs[k] == alone[k][1], (k, sizes[k], rs[k], alone[k][1])
        assert np.array_equal(bits(us[k]), bits(alone[k][0])), (k, sizes[k])
    for p in probs:
        p.close()
    c.close()
