"""The oracle's judge (oracle/oracle.cpp orc_judge, orc_hinge_of, orc_lerp_load,
orc_graph_components; oracle.cascade) pinned without the reference's World, which does not build
on its own:

- phi against the reference's own bond_load + bond_utilization (oracle/_ref) where that builds;
- the picks against a brute-force ordering of those phi (numpy lexsort) and against the closed
  forms of the quota rules (band so high that only the quota is taken; band 0: everything over);
- the DIF lerp by its fixed points (baseline = loads leaves phi as it is) and by homogeneity
  (from a zero baseline, a tension bond's phi scales by dif);
- hinge_of by sections whose answer is fixed by geometry (the pivot at the compression edge, the
  plastic moment hinge_shape x fb x s / c);
- graph_components against scipy's connected components;
- the cascade by what its end must satisfy (nothing over, what left is unsupported).
"""
import numpy as np
import pytest

import oracle
from oracle import CpuJudge, CpuProblem
from opus_destruct_2_amd import gen
from opus_destruct_2_amd.abi import BOND, BOND_LOAD, MATERIAL, SVXG_JUDGE_PIECE, ductile_mask
from opus_destruct_2_amd.stress import default_judge_options


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def _overloaded(mats, shape=(8, 8, 6), voids=0.15, seed=5, target=3.0):
    """A voided block under gravity with strengths scaled so that its largest phi is `target`."""
    g = gen.block(*shape, 0.5, voids, seed=seed)
    p = CpuProblem(g.nodes, g.bonds, mats)
    assert p.assemble()
    u, r = p.solve(g.f, rtol=1e-10, maxit=20000)
    assert r["converged"]
    phimax = p.bond_checks(u)[0].max()
    return g, p, u, target / phimax


def _brute_picks(phi, judged, opts):
    """The round's picks by the reference's rule (world.cpp:1213-1253), ordered by numpy."""
    idx = np.flatnonzero(judged & (phi >= 1.0))
    if len(idx) == 0:
        return idx
    order = idx[np.lexsort((idx, -phi[idx]))]
    n = len(order)
    thr = max(1.0, opts.break_band * phi[order[0]])
    quota = int(np.ceil(min(max(opts.quota_fraction, 0.0), 1.0) * n)) if opts.rule == SVXG_JUDGE_PIECE \
        else max(1, -(-n // 4))
    take = min(n, opts.max_breaks, max(int(np.count_nonzero(phi[order] >= thr)), quota))
    return order[:take]


def test_judge_phi_is_the_reference_bond_checks(mats):
    g, p, u, fr = _overloaded(mats)
    J = CpuJudge(p)
    opts = default_judge_options(fragility=fr)
    j = J.judge(u, opts, {}, want_phi=True)
    kind = "ref" if oracle.ref_available() else "oracle"
    q = CpuProblem(g.nodes, g.bonds, mats, kind=kind)
    phi_ref, _ = q.bond_checks(u, fragility=fr)
    assert np.array_equal(_bits(j["phi"]), _bits(phi_ref))
    assert j["max_phi"] == phi_ref.max() and j["over"] == int(np.count_nonzero(phi_ref >= 1.0))


@pytest.mark.parametrize("rule,frac,band,cap", [(0, 0.25, 0.85, 256), (0, 0.25, 0.85, 3), (1, 0.25, 0.85, 256),
                                                (1, 0.0, 0.95, 256), (1, 1.0, 0.85, 7), (0, 0.25, 0.0, 4096)])
def test_picks_are_the_ordered_band_and_quota(rule, frac, band, cap, mats):
    g, p, u, fr = _overloaded(mats, target=4.0)
    opts = default_judge_options(fragility=fr, rule=rule, quota_fraction=frac, break_band=band, max_breaks=cap)
    j = CpuJudge(p).judge(u, opts, {}, want_phi=True)
    phi = j["phi"]
    want = _brute_picks(phi, np.ones(len(phi), bool), opts)
    assert j["over"] > 8
    assert list(j["picks"]["bond"]) == list(want)
    assert np.array_equal(_bits(j["picks"]["phi"]), _bits(phi[want]))
    assert np.all(j["picks"]["round"] == 0)


def test_quota_closed_forms(mats):
    g, p, u, fr = _overloaded(mats, target=4.0)
    J = CpuJudge(p)
    n = J.judge(u, default_judge_options(fragility=fr), {})["over"]
    assert n >= 9
    # a band no phi reaches: the structure rule takes ceil(over / 4), the piece rule ceil(f x over)
    hi = default_judge_options(fragility=fr, break_band=1e9)
    assert J.judge(u, hi, {})["picked"] == -(-n // 4)
    for frac in (0.1, 0.5, 1.0):
        o = default_judge_options(fragility=fr, break_band=1e9, rule=SVXG_JUDGE_PIECE, quota_fraction=frac)
        assert J.judge(u, o, {})["picked"] == int(np.ceil(frac * n))
    # band 0: everything over strength, up to the cap
    assert J.judge(u, default_judge_options(fragility=fr, break_band=0.0, max_breaks=4096), {})["picked"] == n
    assert J.judge(u, default_judge_options(fragility=fr, break_band=0.0, max_breaks=5), {})["picked"] == 5
    # past max_rounds: judged, nothing picked
    r = J.judge(u, default_judge_options(fragility=fr, max_rounds=3), {"rounds": 3})
    assert r["over"] == n and r["picked"] == 0


def test_ties_go_by_index(mats):
    # every bond given bond 0's section and a vertical stretch u_z = e z: the bonds of a storey
    # carry bitwise equal loads, so phi ties and the order falls to the index
    g = gen.tower(4, 4, 6, seed=2)
    b = g.bonds.copy()
    for k in ("area", "s1", "s2", "c1", "c2", "rmax", "la", "lb", "strength", "sectioned", "ft", "fb", "fc", "coh",
              "mu"):
        b[k] = b[k][0]
    p = CpuProblem(g.nodes, b, mats)
    u = np.zeros(6 * g.n)
    u[2::6] = 1e-3 * g.nodes["c"][:, 2]
    phi, _ = p.bond_checks(u)
    opts = default_judge_options(fragility=3.0 / phi.max(), break_band=0.0, max_breaks=4096)
    j = CpuJudge(p).judge(u, opts, {}, want_phi=True)
    ph, bd = j["picks"]["phi"], j["picks"]["bond"]
    assert len(ph) > 20
    assert np.all((ph[:-1] > ph[1:]) | ((ph[:-1] == ph[1:]) & (bd[:-1] < bd[1:])))
    assert np.count_nonzero(ph[:-1] == ph[1:]) > 10  # (the case is exercised)
    want = _brute_picks(j["phi"], np.ones(len(b), bool), opts)
    assert list(bd) == list(want)


def test_dif_lerp_fixed_points_and_scaling(mats):
    g, p, u, fr = _overloaded(mats)
    J = CpuJudge(p)
    opts = default_judge_options(fragility=fr, dif=1.5)
    static = J.judge(u, opts, {}, want_phi=True)
    loads = p.bond_loads(u)
    # the judgement left its loads as the baseline
    assert J.has.all()
    assert np.array_equal(J.base.view(np.uint64), loads.view(np.uint64))
    # a shock measured from those same loads changes nothing
    same = J.judge(u, opts, {"shock": True}, want_phi=True)
    assert np.array_equal(_bits(same["phi"]), _bits(static["phi"]))
    # from a zero baseline the shock's loads are dif x L: tension bonds' phi is homogeneous in L
    J.base[:] = 0
    shocked = J.judge(u, opts, {"shock": True, "transient": True}, want_phi=True)
    ten = loads["N"] > 0
    assert ten.sum() > 10
    np.testing.assert_allclose(shocked["phi"][ten], 1.5 * static["phi"][ten], rtol=1e-12)
    # transient: the baseline stayed zero; no shock: no lerp
    assert np.all(J.base.view(np.uint64) == 0)
    calm = J.judge(u, opts, {"shock": False}, want_phi=True)
    assert np.array_equal(_bits(calm["phi"]), _bits(static["phi"]))
    # bonds without a baseline are judged statically under a shock
    J.has[:] = 0
    J.base[:] = 0
    fresh = J.judge(u, opts, {"shock": True}, want_phi=True)
    assert np.array_equal(_bits(fresh["phi"]), _bits(static["phi"]))


def test_lerp_load_endpoints():
    rng = np.random.default_rng(3)
    a = np.zeros(1, BOND_LOAD)
    b = np.zeros(1, BOND_LOAD)
    for k in BOND_LOAD.names:
        a[k], b[k] = rng.normal(size=2) * 1e4
    assert oracle.lerp_load(a, b, 0.0).tobytes() == a[0].tobytes()
    one, two = oracle.lerp_load(a, b, 1.0), oracle.lerp_load(a, b, 2.0)
    for k in BOND_LOAD.names:
        assert np.isclose(one[k], b[0][k], rtol=1e-14, atol=1e-10)
        assert np.isclose(two[k], 2 * b[0][k] - a[0][k], rtol=1e-14, atol=1e-10)
    z = np.zeros(1, BOND_LOAD)
    r = oracle.lerp_load(z, b, 2.5)
    for k in BOND_LOAD.names:
        assert r[k] == 2.5 * b[0][k]


def _section(sectioned=True):
    """A bond along z (n = z, t1 = x, t2 = y): area 0.04, s1 = s2 = 1e-4, c1 = c2 = 0.1."""
    b = np.zeros(1, BOND)
    b["a"], b["b"] = 0, 1
    b["n"], b["t1"], b["t2"] = [0, 0, 1], [1, 0, 0], [0, 1, 0]
    b["area"], b["p"] = 0.04, [1.0, 2.0, 3.0]
    b["s1"], b["s2"], b["c1"], b["c2"], b["rmax"] = 1e-4, 2e-4, 0.1, 0.2, 0.15
    b["la"], b["lb"], b["strength"] = 0.25, 0.25, 1.0
    b["sectioned"] = 1 if sectioned else 0
    b["ft"], b["fb"], b["fc"], b["coh"], b["mu"] = 2e6, 3e6, 2e7, 1e6, 0.6
    return b


def test_hinge_known_answers():
    mats = np.zeros(2, MATERIAL)
    mats["E"], mats["G"] = 3e10, 1.2e10
    duct = ductile_mask([0])
    b = _section()
    M1 = 5e3
    load = {"M1": M1}
    # the far side turning about t1: the axis is t1, the pivot the compression edge along t2
    # (cross(t1, n) = -t2: p + 0.9 c2 t2), the plastic moment 1.3 fb s2 / c2, the pull ft A
    h = oracle.hinge_of(b, load, [0, 0, 0], [1e-3, 0, 0], mats, duct, hinge_shape=1.3)
    assert h is not None
    assert np.allclose(h["axis"], [1, 0, 0]) and np.allclose(h["n"], [0, 0, 1])
    assert np.allclose(h["p"], [1.0, 2.0 + 0.9 * 0.2, 3.0])
    assert np.isclose(h["mp"], 1.3 * 3e6 * 2e-4 / 0.2) and np.isclose(h["pull"], 2e6 * 0.04)
    assert h["a"] == 0 and h["b"] == 1
    # a rotation along n only (no turn square to the bond): the axis follows the moment, the pivot
    # stays at the centroid
    h = oracle.hinge_of(b, load, [0, 0, 0], [0, 0, 1e-3], mats, duct)
    assert np.allclose(h["axis"], [1, 0, 0]) and np.allclose(h["p"], [1.0, 2.0, 3.0])
    # turning about t2 the other way: axis -t2, cross(-t2, n) = -t1, so the pivot is p + 0.9 c1 t1
    h = oracle.hinge_of(b, {"M2": 4e3}, [0, 1e-3, 0], [0, 0, 0], mats, duct)
    assert np.allclose(h["axis"], [0, -1, 0]) and np.allclose(h["p"], [1.0 + 0.9 * 0.1, 2.0, 3.0])
    assert np.isclose(h["mp"], 1.3 * 3e6 * 1e-4 / 0.1)
    # a support does not turn: only a's rotation counts
    s = b.copy()
    s["b"] = -1
    h = oracle.hinge_of(s, load, [-1e-3, 0, 0], [5, 5, 5], mats, duct)
    assert np.allclose(h["axis"], [1, 0, 0])
    # not ductile on both sides, or pull / shear governing: no hinge
    m2 = b.copy()
    m2["mb"] = 1
    assert oracle.hinge_of(m2, load, [0, 0, 0], [1e-3, 0, 0], mats, duct) is None
    assert oracle.hinge_of(b, {"M1": M1, "N": 0.9 * 0.04 * 2e6 * 10}, [0, 0, 0], [1e-3, 0, 0], mats, duct) is None
    assert oracle.hinge_of(b, {"M1": 1.0, "V1": 1e6}, [0, 0, 0], [1e-3, 0, 0], mats, duct) is None
    assert oracle.hinge_of(b, {}, [0, 0, 0], [0, 0, 0], mats, duct) is None


def test_graph_components_match_scipy():
    sp = pytest.importorskip("scipy.sparse")
    csg = pytest.importorskip("scipy.sparse.csgraph")
    rng = np.random.default_rng(11)
    n, m = 300, 420
    b = np.zeros(m, BOND)
    b["a"] = rng.integers(0, n, m)
    b["b"] = rng.integers(-1, n, m)
    b["broken"] = rng.random(m) < 0.2
    seed = np.zeros(n, np.uint8)
    seed[rng.integers(0, n, 6)] = 1
    nc, comp = oracle.graph_components(n, b, seed)
    keep = (b["broken"] == 0) & (b["b"] >= 0)
    A = sp.coo_matrix((np.ones(keep.sum()), (b["a"][keep], b["b"][keep])), shape=(n, n))
    k, lab = csg.connected_components(A, directed=False)
    # the same partition; component 0 = every component holding a seed
    for c in range(k):
        ours = np.unique(comp[lab == c])
        assert len(ours) == 1
    seeded = np.isin(lab, np.unique(lab[seed == 1]))
    assert np.array_equal(comp == 0, seeded)
    assert nc == 1 + len(np.unique(lab[~seeded]))
    # the rest numbered in order of their first node
    firsts = [int(np.flatnonzero(comp == c)[0]) for c in range(1, nc)]
    assert firsts == sorted(firsts)


def test_cascade_end_state(mats):
    g = gen.block(10, 10, 8, 0.5, 0.3, seed=5)
    p = CpuProblem(g.nodes, g.bonds, mats)
    assert p.assemble()
    u0, _ = p.solve(g.f, rtol=1e-10, maxit=20000)
    fr = 3.0 / p.bond_checks(u0)[0].max()
    q = CpuProblem(g.nodes, g.bonds, mats)
    J = CpuJudge(q)
    opts = default_judge_options(fragility=fr)
    u, res = oracle.cascade(q, J, g.f, np.zeros(6 * g.n), opts, {})
    assert res["end"] == "settled" and res["rounds"] >= 3 and len(res["breaks"]) >= res["rounds"]
    rnd = res["breaks"]["round"]
    assert np.all(np.diff(rnd) >= 0) and set(rnd) == set(range(res["rounds"]))
    # nothing over strength in the final state (a judgement without shock)
    assert J.judge(u, opts, {"shock": False})["over"] == 0
    # every broken bond broke in a round or with a retired node; the retired have no supported path
    bonds = q.get_bonds()
    nodes = q.get_nodes()
    gone = set(np.flatnonzero(nodes["gone"]))
    assert gone == set(res["retired"])
    for k in np.flatnonzero(bonds["broken"]):
        assert k in set(res["breaks"]["bond"]) or bonds["a"][k] in gone or bonds["b"][k] in gone
    seed = np.zeros(g.n, np.uint8)
    sup = (bonds["broken"] == 0) & (bonds["b"] < 0)
    seed[bonds["a"][sup]] = 1
    _, comp = oracle.graph_components(g.n, bonds, seed)
    assert np.all(comp[nodes["gone"] == 0] == 0)
    assert np.all(u.reshape(-1, 6)[nodes["gone"] == 1] == 0)
