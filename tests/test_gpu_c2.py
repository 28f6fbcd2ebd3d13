"""Parity on the single-large-structure workload itself: the 64 x 64 x 48 block with 15 % voids and
materials mixed per column (167,129 nodes, 1.0M DOFs), on the settings a graph of this size takes by
default (mixed precision: the V-cycle over FP16 block copies and the lean hierarchy), on the mixed
split with FP32 K products inside FP64 refinement, and on the plain FP64 path, each solved to a
relative residual of 1e-10 and compared with the reference's own CPU solve of the same block
(oracle/_ref, compiled from the reference's sources; the oracle's restatement where _ref is absent,
bit-identical to it on every case of test_oracle).

Bars: u within 1e-6 max|u_ref| (FP64) / 1e-5 (mixed); phi within 1e-6 max(1, |phi_ref|) (FP64) /
1e-5 (mixed); the failed bonds {phi >= 1} identical except bonds within 1e-4 of phi = 1.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import CpuProblem, reference_options
from opus_destruct_2_amd import gen, stress
from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED
from tests.test_gpu_parity import failed_sets_agree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def block(mats):
    g = gen.block(64, 64, 48, 0.5, 0.15, seed=1)
    kind = "ref" if oracle.ref_available() else "oracle"
    if kind == "ref":
        oracle.set_ref_threads(min(16, len(os.sched_getaffinity(0))))
    o = CpuProblem(g.nodes, g.bonds, mats, kind)
    assert o.assemble(reference_options())
    uo, ro = o.solve(g.f, rtol=1e-10, maxit=20000)
    assert ro["converged"] and ro["rel_res"] <= 1e-10
    phi_o, _ = o.bond_checks(uo)
    print(f"c2 block: {g.n} nodes, {g.m} bonds; CPU solve ({kind}) {ro['iters']} iterations, "
          f"{int((phi_o >= 1).sum())} bonds failed", flush=True)
    return g, uo, phi_o


PATHS = [("mixed", SVXG_MIXED, {}, 1e-5), ("mixed_fp32_products_ir", SVXG_MIXED, {"mixed_fp64_rows": 0}, 1e-5),
         ("fp64", SVXG_FP64, {}, 1e-6)]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("path", PATHS, ids=[p[0] for p in PATHS])
def test_block_matches_reference(path, block, mats, ctx):
    name, precision, kw, tol = path
    g, uo, phi_o = block
    d = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    assert d.assemble(stress.default_options(precision=precision, **kw))
    info = d.info()
    # (the defaults of a graph of 16384+ free nodes: the lean hierarchy ending in a dense level, and
    # on the mixed path the V-cycle over FP16 block copies)
    assert info["lean"] and info["dense_rows"] > 0, info
    assert info["vcycle_fp16"] == (precision == SVXG_MIXED), info
    ud, rd = d.solve(g.f, rtol=1e-10, maxit=20000)
    assert rd["converged"] and rd["rel_res"] <= 1e-10, rd
    scale = np.abs(uo).max()
    u_err = np.abs(ud - uo).max() / scale
    phi_d, _ = d.bond_checks(None)
    phi_err = np.abs(phi_d - phi_o) / np.maximum(1.0, np.abs(phi_o))
    print(f"c2 {name}: {rd['iters']} iterations, u error {u_err:.2e}, phi error {phi_err.max():.2e}, "
          f"failed {int((phi_d >= 1).sum())} vs {int((phi_o >= 1).sum())}", flush=True)
    assert u_err <= tol, u_err
    assert phi_err.max() <= tol, phi_err.max()
    assert failed_sets_agree(phi_d, phi_o)
    d.close()
