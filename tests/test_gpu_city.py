"""The city chunk at test size: resident buildings solved in one batch under changing loads; a
ground-storey collapse (the lowest storey's columns cut under half the footprint) removes bonds in
place and triggers a re-setup of that building's multigrid. Each answer matches the oracle and the
device's FP64 solve (1e-5 of max |u|; failed bonds equal outside the 1e-4 band), and a collapsed,
re-set-up building gives bitwise the answer of a fresh problem built with those bonds broken."""
import numpy as np
import pytest

from oracle import CpuProblem, reference_options
from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED

pytestmark = pytest.mark.gpu

SHAPES = [(6, 7, 12), (8, 8, 16), (7, 8, 24), (9, 9, 14), (5, 6, 30)]


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def failed_sets_agree(phi, phi_ref, band=1e-4):
    diff = (phi >= 1.0) != (phi_ref >= 1.0)
    near = (np.abs(phi - 1.0) <= band) | (np.abs(phi_ref - 1.0) <= band)
    return not np.any(diff & ~near)


def test_collapse_resetup_and_sample_checks(mats, ctx):
    gs = [gen.building(*s, seed=40 + k) for k, s in enumerate(SHAPES)]
    opts = stress.default_options(precision=SVXG_MIXED)
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    assert all(stress.assemble_batch(probs))
    # round 0: wind from +x
    us, rs = stress.solve_batch(probs, [gen.wind_load(g, 0.0) for g in gs], rtol=1e-8)
    assert all(r["converged"] for r in rs), rs
    # collapses of buildings 1 and 3: bonds out in place, then a re-setup of those two only
    hit = [1, 3]
    cuts = {k: gen.ground_storey_cuts(gs[k]) for k in hit}
    for k in hit:
        before = probs[k].info()["blocks"]
        probs[k].remove_bonds(cuts[k])
        assert probs[k].bonds["broken"][cuts[k]].all()
        assert probs[k].info()["blocks"] == before  # (pattern kept until the re-setup)
    assert all(stress.assemble_batch([probs[k] for k in hit]))
    # round 1: wind from +y, all buildings
    fs = [gen.wind_load(g, np.pi / 2) for g in gs]
    us, rs = stress.solve_batch(probs, fs, rtol=1e-10, maxit=20000)
    assert all(r["converged"] for r in rs), rs
    # a fresh problem built with the collapse's bonds broken gives the same bits
    fresh = []
    for k in hit:
        b = gs[k].bonds.copy()
        b["broken"][cuts[k]] = 1
        fresh.append(stress.StressProblem(gs[k].nodes, b, mats, ctx, options=opts))
    assert all(stress.assemble_batch(fresh))
    uf, rf = stress.solve_batch(fresh, [fs[k] for k in hit], rtol=1e-10, maxit=20000)
    for j, k in enumerate(hit):
        assert rf[j] == rs[k], (k, rf[j], rs[k])
        assert np.array_equal(bits(uf[j]), bits(us[k])), k
    # every building against the oracle and against the device's FP64 path
    for k, g in enumerate(gs):
        b = g.bonds.copy()
        if k in cuts:
            b["broken"][cuts[k]] = 1
        o = CpuProblem(g.nodes, b, mats)
        assert o.assemble(reference_options())
        uo, ro = o.solve(fs[k], rtol=1e-10, maxit=20000)
        assert ro["converged"]
        scale = np.abs(uo).max()
        assert np.abs(us[k] - uo).max() <= 1e-5 * scale, (k, np.abs(us[k] - uo).max() / scale)
        p64 = stress.StressProblem(g.nodes, b, mats, ctx, options=stress.default_options(precision=SVXG_FP64))
        assert p64.assemble()
        u64, r64 = p64.solve(fs[k], rtol=1e-10, maxit=20000)
        assert r64["converged"]
        assert np.abs(u64 - uo).max() <= 1e-6 * scale
        phi, _ = probs[k].bond_checks(us[k])
        phi64, _ = p64.bond_checks(u64)
        phio, _ = o.bond_checks(uo)
        assert failed_sets_agree(phi, phi64) and failed_sets_agree(phi, phio), k
        assert np.abs(phi - phio).max() <= 1e-5 * max(1.0, np.abs(phio).max()), k
        p64.close()
    for p in probs + fresh:
        p.close()
