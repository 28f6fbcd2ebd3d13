"""Writes the golden fixtures tests/golden/*.npz from the reference's own stress path compiled
from its sources (oracle/_ref/libsvxref.so, built by oracle/Makefile where /root/reference is).

Each fixture holds a problem (nodes, bonds, materials as raw struct bytes; the load; the options),
and the reference's answer: u (all nodes), the iteration count and convergence, and every bond's
utilisation phi and failure mode. Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import CpuProblem, reference_options  # noqa: E402
from opus_destruct_2_amd import gen  # noqa: E402
from tests import problems as P  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
KEYS = ["rtol", "amg_min_nodes", "strength", "coarse_scale", "coarse_max", "smoothed", "max_levels", "sweeps",
        "coarsen_dense", "reaggregate"]


def opt_pairs(**kw):
    o = reference_options(**kw)
    vals = [o.rtol, o.amg_min_nodes, o.amg.strength, o.amg.coarse_scale, o.amg.coarse_max, o.amg.smoothed,
            o.amg.max_levels, o.amg.sweeps, o.amg.coarsen_dense, o.amg.reaggregate]
    return o, np.array(KEYS), np.array(vals, np.float64)


def write(name, nodes, bonds, f, rtol, maxit, **kw):
    mats = gen.materials()
    o, keys, vals = opt_pairs(**kw)
    p = CpuProblem(nodes, bonds, mats, "ref")
    assert p.assemble(o)
    u, r = p.solve(f, rtol=rtol, maxit=maxit)
    phi, mode = p.bond_checks(u)
    np.savez_compressed(os.path.join(OUT, name + ".npz"),
                        nodes=np.frombuffer(nodes.tobytes(), np.uint8), bonds=np.frombuffer(bonds.tobytes(), np.uint8),
                        mats=np.frombuffer(mats.tobytes(), np.uint8), f=f, opt_keys=keys, opt_vals=vals,
                        rtol=rtol, maxit=maxit, u=u, iters=r["iters"], converged=r["converged"], phi=phi, mode=mode)
    print(name, r, len(nodes), len(bonds))


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libsvxref.so is not built (make -C oracle)")
    nodes, bonds = P.cantilever(12, 0.5, 300.0)
    write("cantilever12", nodes, bonds, P.gravity(nodes), 1e-10, 500, amg_min_nodes=4)
    nodes, bonds = P.cube_block(8, 0.4)
    write("block8_amg", nodes, bonds, P.gravity(nodes), 1e-9, 2000, amg_min_nodes=16)
    nodes, bonds = P.cube_block(10, 0.4, 7, True)
    write("block10_retired", nodes, bonds, P.gravity_live(nodes), 1e-8, 500, amg_min_nodes=16)
    g = gen.tower()
    write("tower8x8x20", g.nodes, g.bonds, g.f, 1e-8, 5000)
    g = gen.piece(400, pinned=True, seed=7)
    write("piece400_pinned", g.nodes, g.bonds, g.f, 1e-8, 5000, amg_min_nodes=0)
    g = gen.piece(120, pinned=False, seed=8)
    write("piece120_jacobi", g.nodes, g.bonds, g.f, 1e-8, 5000)


if __name__ == "__main__":
    main()
