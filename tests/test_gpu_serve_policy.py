"""The serving loop's failure policy (step_structures, core/src/world/world.cpp:1056-1100) and its
judgement of converged requests (world.cpp:1111), with the failures forced from the host: a load
with non-finite values breaks a run down (the policy zeroes them, rebuilds, detaches what lost its
supports - world.cpp:1371-1380 - and restarts), tiny run budgets exhaust solve_restarts (the solve
is abandoned), and a judged request's picks equal svxg_judge_batch's on the same solution."""
import numpy as np
import pytest

from oracle import CpuProblem, reference_options
from opus_destruct_2_amd import gen, serve, stress
from opus_destruct_2_amd.abi import SVXG_MIXED, SVXG_OK

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_breakdown_restarts_on_the_finite_load(mats, ctx):
    g = gen.piece(400, pinned=True, seed=31)
    f = g.f.copy()
    bad = [6 * 7 + 2, 6 * 50 + 0, 6 * 51 + 4]
    f[bad[0]], f[bad[1]], f[bad[2]] = np.nan, np.inf, -np.inf
    srv = serve.Server(ctx, precision=SVXG_MIXED, rtol=1e-9)
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    t = srv.submit(p, f)
    rec = srv.wait([t])[t]
    st = srv.stats()
    srv.close()
    assert rec["status"] == SVXG_OK and rec["result"]["converged"], rec
    assert rec["breakdowns"] == 1 and rec["restarts"] == 1 and rec["detached"] == 0, rec
    assert st["breakdowns"] == 1 and st["solves_abandoned"] == 0
    # the same as the problem's own solve of the load with those values zeroed
    f0 = f.copy()
    f0[~np.isfinite(f0)] = 0.0
    q = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=stress.default_options(precision=SVXG_MIXED))
    assert q.assemble()
    u0, r0 = q.solve(f0, rtol=1e-9)
    assert r0["converged"]
    assert np.abs(rec["u"] - u0).max() <= 1e-9 * np.abs(u0).max()
    q.close()


def loose_cluster(g, seed_node):
    """A node and its bonded neighbours, cut from the rest and from their supports: (cluster, cuts)."""
    b = g.bonds
    cl = {seed_node}
    for k in range(len(b)):
        if b["b"][k] >= 0 and (b["a"][k] == seed_node or b["b"][k] == seed_node):
            cl.add(int(b["a"][k]))
            cl.add(int(b["b"][k]))
    cuts = [k for k in range(len(b))
            if (b["b"][k] < 0 and int(b["a"][k]) in cl)
            or (b["b"][k] >= 0 and ((int(b["a"][k]) in cl) != (int(b["b"][k]) in cl)))]
    return sorted(cl), cuts


def loose_after(g, cuts):
    """The nodes with no path over intact bonds to a support bond once `cuts` broke (BFS)."""
    b = g.bonds
    cut = np.zeros(len(b), bool)
    cut[cuts] = True
    adj = [[] for _ in range(g.n)]
    reached = np.zeros(g.n, bool)
    stack = []
    for k in range(len(b)):
        if cut[k]:
            continue
        a, c = int(b["a"][k]), int(b["b"][k])
        if c < 0:
            if not reached[a]:
                reached[a] = True
                stack.append(a)
        else:
            adj[a].append(c)
            adj[c].append(a)
    while stack:
        k = stack.pop()
        for j in adj[k]:
            if not reached[j]:
                reached[j] = True
                stack.append(j)
    return sorted(np.nonzero(~reached)[0].tolist())


def test_breakdown_detaches_what_lost_its_supports(mats, ctx):
    g = gen.piece(500, pinned=False, seed=12)  # (on supports)
    top = int(np.argmax(g.nodes["c"][:, 2]))
    cluster, cuts = loose_cluster(g, top)
    assert 1 < len(cluster) < g.n // 4
    f = g.f.copy()
    f[6 * cluster[0] + 2] = np.nan  # (forces the breakdown that starts the policy)
    srv = serve.Server(ctx, precision=SVXG_MIXED, rtol=1e-10)
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    t = srv.submit(p, f, remove=cuts)
    rec = srv.wait([t])[t]
    srv.close()
    assert rec["result"]["converged"] and rec["breakdowns"] >= 1, rec
    loose = loose_after(g, cuts)
    assert set(cluster) <= set(loose)
    assert rec["detached"] == len(loose), (rec, len(loose))
    nodes = p.nodes
    assert all(nodes["gone"][loose])
    cluster = loose
    # the oracle on the same graph: those bonds out, the cluster retired, the load made finite
    f0 = f.copy()
    f0[~np.isfinite(f0)] = 0.0
    o = CpuProblem(g.nodes, g.bonds, mats)
    for k in cuts:
        o.set_bond_broken(k, 1)
    o.retire_nodes(cluster)
    assert o.assemble(reference_options())
    uo, ro = o.solve(f0, rtol=1e-10, maxit=20000)
    assert ro["converged"]
    assert np.abs(rec["u"] - uo).max() <= 1e-5 * np.abs(uo).max()
    assert np.all(rec["u"].reshape(-1, 6)[cluster] == 0.0)


def test_abandoned_after_solve_restarts(mats, ctx):
    g = gen.piece(600, pinned=False, seed=5)
    srv = serve.Server(ctx, precision=SVXG_MIXED, rtol=1e-10, restart_iters=3, solve_restarts=2)
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    t = srv.submit(p, g.f)
    rec = srv.wait([t])[t]
    st = srv.stats()
    srv.close()
    assert rec["abandoned"] and not rec["result"]["converged"], rec
    assert rec["restarts"] == 2 and rec["result"]["iters"] == 6, rec
    assert st["solves_abandoned"] == 1 and st["restarts"] == 1
    assert np.all(np.isfinite(rec["u"]))


def test_judged_requests_pick_as_the_judge(mats, ctx):
    g = gen.piece(800, pinned=False, seed=44)
    probe = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=stress.default_options(precision=SVXG_MIXED))
    assert probe.assemble()
    u, r = probe.solve(g.f, rtol=1e-9)
    phi, _ = probe.bond_checks(None)
    jo = stress.default_judge_options(fragility=2.0 / float(phi.max()))
    ref = stress.judge_batch([probe], [{}], jo)[0]
    probe.close()
    assert ref["picked"] > 0
    srv = serve.Server(ctx, precision=SVXG_MIXED, rtol=1e-9, judge=jo)
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    t = srv.submit(p, g.f)
    rec = srv.wait([t])[t]
    picks = srv.picks(p)
    st = srv.stats()
    srv.close()
    assert rec["result"]["converged"] and st["judged"] == 1
    assert rec["judge"]["over"] == ref["over"] and rec["judge"]["picked"] == ref["picked"], (rec["judge"], ref)
    assert rec["judge"]["max_phi"] == ref["max_phi"]
    assert np.array_equal(picks["bond"], ref["picks"]["bond"])
