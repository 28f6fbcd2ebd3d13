"""The dense coarsest level (amg.cpp:535-592: a Cholesky of the coarsest matrix, here turned into
an explicit inverse on the matrix cores). A graph that never coarsens is a single dense level,
so its preconditioner is K^-1 itself: applied to r it must match a direct solve of the assembled
K, and PCG must converge in one iteration. Sizes cover one 16-row tile, a padded tile edge, and
the largest dense level (192 nodes, 1152 rows)."""
import numpy as np
import pytest

from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED

pytestmark = pytest.mark.gpu


def dense_k(p):
    rowptr, col, val = p.matrix()
    n = len(rowptr) - 1
    K = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for k in range(rowptr[i], rowptr[i + 1]):
            j = col[k]
            K[6 * i:6 * i + 6, 6 * j:6 * j + 6] = val[k]
    return K


@pytest.mark.parametrize("count,pinned,coarse_max", [(2, False, 48), (3, True, 48), (31, False, 48),
                                                      (48, True, 48), (120, False, 192), (192, False, 192)])
@pytest.mark.parametrize("precision", [SVXG_FP64, SVXG_MIXED], ids=["fp64", "mixed"])
def test_single_dense_level_is_the_exact_inverse(count, pinned, coarse_max, precision, mats, ctx):
    g = gen.piece(count, pinned=pinned, seed=600 + count)
    opts = stress.default_options(precision=precision, amg_min_nodes=0, coarse_max=coarse_max)
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts)
    assert p.assemble(opts)
    info = p.info()
    assert info["levels"] == 1 and info["level_rows"][0] == g.nfree
    K = dense_k(p)
    r = gen.uniform(31 + count, K.shape[0]) - 0.5
    z = p.apply_preconditioner(r)
    # (mixed: the V-cycle takes r rounded to FP32 and returns z in FP32)
    rin = r if precision == SVXG_FP64 else r.astype(np.float32).astype(np.float64)
    ref = np.linalg.solve(K, rin)
    err = np.abs(z - ref).max() / np.abs(ref).max()
    assert err <= (1e-8 if precision == SVXG_FP64 else 1e-6), err
    u, res = p.solve(g.f, rtol=1e-8, maxit=50)
    assert res["converged"], res
    if precision == SVXG_FP64:
        assert res["iters"] <= 2, res
    p.close()


def test_dense_levels_of_many_sizes_in_one_setup(mats, ctx):
    """Dense levels of one tile row up to 72 (the one-workgroup factorisation up to 18 tile rows, the
    many-workgroup one above), set up in one batch: each the exact inverse, and each bit for bit the
    same as set up alone."""
    sizes = [3, 31, 48, 49, 80, 120, 192]
    opts = stress.default_options(precision=SVXG_FP64, amg_min_nodes=0, coarse_max=192)
    gs = [gen.piece(n, pinned=bool(k & 1), seed=700 + n) for k, n in enumerate(sizes)]
    batch = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    stress.assemble_batch(batch, ctx)
    for g, p in zip(gs, batch):
        assert p.info()["levels"] == 1
        K = dense_k(p)
        r = gen.uniform(77, K.shape[0]) - 0.5
        z = p.apply_preconditioner(r)
        ref = np.linalg.solve(K, r)
        assert np.abs(z - ref).max() <= 1e-8 * np.abs(ref).max(), g.n
        alone = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts)
        assert alone.assemble(opts)
        assert np.array_equal(alone.apply_preconditioner(r).view(np.uint64), z.view(np.uint64)), g.n
        alone.close()
        p.close()
