"""The autotuner: tuning a level shape online - launches of its candidate tilings taken in turn,
timed, the fastest kept - changes no bit of any answer (a row's sums do not depend on the tile that
holds it), and a context records its choices. Solves on a fresh context that tunes against the same
solves on one that does not (the untuned defaults): the single large block's generator at a small
size with FP16 copies and the lean hierarchy forced (every tuned launch kind: smoothing,
prolongation, FP16 restriction, dense solve), the same on the FP64 path, and a batch of pieces."""
import numpy as np
import pytest

from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def solve_all(ctx, gs, mats, tune, **kw):
    ctx.set_autotune(tune)
    opts = stress.default_options(**kw)
    ps = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    assert all(stress.assemble_batch(ps, ctx=ctx))
    out = []
    for _ in range(2):  # (the second after the first has tuned)
        us, rs = stress.solve_batch(ps, [g.f for g in gs], rtol=1e-10, maxit=5000, ctx=ctx)
        assert all(r["converged"] for r in rs)
        out.append((us, rs))
    for p in ps:
        p.close()
    return out


@pytest.mark.parametrize("kind", ["block_mixed_fp16_lean", "block_fp64", "pieces_mixed"])
def test_tuned_bits_equal_untuned(kind, mats):
    if kind == "pieces_mixed":
        gs = [gen.piece(int(n), pinned=bool(k & 1), seed=700 + k) for k, n in enumerate(gen.piece_sizes(48, seed=9))]
        kw = dict(precision=SVXG_MIXED, amg_min_nodes=0)
    else:
        gs = [gen.block(20, 20, 16, 0.5, 0.15, seed=4)]
        kw = (dict(precision=SVXG_MIXED, vcycle_fp16_rows=0, lean_rows=1) if kind == "block_mixed_fp16_lean"
              else dict(precision=SVXG_FP64))
    plain, tuned = stress.Context(0), stress.Context(0)
    a = solve_all(plain, gs, mats, False, **kw)
    b = solve_all(tuned, gs, mats, True, **kw)
    rec = tuned.tuned()
    print(kind, rec)
    assert plain.tuned() == []
    assert len(rec) > 0
    assert {"smoothing", "prolongation"} <= {r["kind"] for r in rec}
    if kind == "block_mixed_fp16_lean":
        assert {"restriction", "dense"} <= {r["kind"] for r in rec}
    for (ua, ra), (ub, rb) in zip(a, b):
        assert ra == rb
        for x, y in zip(ua, ub):
            assert np.array_equal(bits(x), bits(y))
    plain.close()
    tuned.close()
