"""numpy / ctypes mirrors of the plain structs in include/svxg.h.

The graph is handed to the C ABI as numpy structured arrays whose layout is the C struct's
(svxg_node 40 B, svxg_bond 216 B, svxg_material 64 B), so a problem is built without a copy on the
Python side. Field names follow the reference's SNode / SBond / Material
(core/include/svx/stress/stress.hpp:21-51, core/include/svx/material/material.hpp:47-57).
"""
import ctypes

import numpy as np

NODE = np.dtype([
    ("c", "<f8", (3,)), ("mass", "<f8"), ("fixed", "u1"), ("gone", "u1"), ("pad_", "u1", (6,)),
], align=True)

BOND = np.dtype([
    ("a", "<i4"), ("b", "<i4"),
    ("broken", "u1"), ("in_k", "u1"), ("ma", "u1"), ("mb", "u1"), ("sectioned", "u1"), ("pad0_", "u1", (3,)),
    ("n", "<f8", (3,)), ("t1", "<f8", (3,)), ("t2", "<f8", (3,)),
    ("area", "<f8"), ("p", "<f8", (3,)),
    ("s1", "<f8"), ("s2", "<f8"), ("c1", "<f8"), ("c2", "<f8"), ("rmax", "<f8"),
    ("la", "<f8"), ("lb", "<f8"), ("strength", "<f8"),
    ("ft", "<f4"), ("fb", "<f4"), ("fc", "<f4"), ("coh", "<f4"), ("mu", "<f4"),
    ("faces", "<i4"), ("tag", "<i4"), ("pad1_", "<i4"),
], align=True)

MATERIAL = np.dtype([
    ("E", "<f8"), ("G", "<f8"), ("rho", "<f8"),
    ("ft", "<f8"), ("fb", "<f8"), ("fc", "<f8"), ("cohesion", "<f8"), ("friction", "<f8"),
], align=True)

BOND_LOAD = np.dtype([("N", "<f8"), ("V1", "<f8"), ("V2", "<f8"), ("T", "<f8"), ("M1", "<f8"), ("M2", "<f8")])

assert NODE.itemsize == 40 and BOND.itemsize == 216 and MATERIAL.itemsize == 64

SVXG_OK, SVXG_BAD_INPUT, SVXG_OOM_BUDGET, SVXG_DEVICE_ERROR, SVXG_NOT_ASSEMBLED, SVXG_SINGULAR = range(6)
SVXG_FP64, SVXG_MIXED = 0, 1


class AmgOptions(ctypes.Structure):
    """AmgOptions (core/include/svx/solve/amg.hpp:46-69)."""
    _fields_ = [
        ("strength", ctypes.c_double), ("coarse_scale", ctypes.c_double),
        ("coarse_max", ctypes.c_int32), ("smoothed", ctypes.c_int32), ("max_levels", ctypes.c_int32),
        ("sweeps", ctypes.c_int32), ("cheb_degree", ctypes.c_int32), ("coarsen_dense", ctypes.c_int32),
        ("reaggregate", ctypes.c_int32), ("smooth_levels", ctypes.c_int32), ("coarse_cheb_degree", ctypes.c_int32),
        ("lean_rows", ctypes.c_int32), ("lean_strength", ctypes.c_double),
    ]


class StressOptions(ctypes.Structure):
    """StressOptions (core/include/svx/stress/stress.hpp:68-72) plus the solve's precision."""
    _fields_ = [("rtol", ctypes.c_double), ("amg_min_nodes", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("amg", AmgOptions), ("mixed_fp64_rows", ctypes.c_int32), ("vcycle_fp16_rows", ctypes.c_int32)]


class PcgResult(ctypes.Structure):
    """PcgResult (core/include/svx/solve/amg.hpp:120-125)."""
    _fields_ = [("iters", ctypes.c_int32), ("converged", ctypes.c_int32), ("breakdown", ctypes.c_int32),
                ("refinements", ctypes.c_int32), ("rel_res", ctypes.c_double)]

    def as_dict(self):
        return {"iters": self.iters, "converged": bool(self.converged), "breakdown": bool(self.breakdown),
                "refinements": self.refinements, "rel_res": self.rel_res}


class ProblemInfo(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int32), ("bonds", ctypes.c_int32), ("nfree", ctypes.c_int32), ("levels", ctypes.c_int32),
        ("block_jacobi", ctypes.c_int32), ("assembled", ctypes.c_int32), ("preconditioner_current", ctypes.c_int32),
        ("vcycle_fp16", ctypes.c_int32),
        ("blocks", ctypes.c_int64), ("work_per_iteration", ctypes.c_int64), ("assembly_work", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("level_rows", ctypes.c_int32 * 16), ("level_blocks", ctypes.c_int64 * 16), ("level_pblocks", ctypes.c_int64 * 16),
        ("lean", ctypes.c_int32), ("dense_rows", ctypes.c_int32),
    ]

    def as_dict(self):
        L = self.levels
        return {
            "nodes": self.nodes, "bonds": self.bonds, "nfree": self.nfree, "levels": L,
            "block_jacobi": bool(self.block_jacobi), "assembled": bool(self.assembled),
            "preconditioner_current": bool(self.preconditioner_current), "vcycle_fp16": bool(self.vcycle_fp16),
            "blocks": self.blocks,
            "work_per_iteration": self.work_per_iteration, "assembly_work": self.assembly_work,
            "device_bytes": self.device_bytes,
            "level_rows": list(self.level_rows[:L]), "level_blocks": list(self.level_blocks[:L]),
            "level_pblocks": list(self.level_pblocks[:L]),
            "lean": bool(self.lean), "dense_rows": self.dense_rows,
        }


class TuneRecord(ctypes.Structure):
    """svxg_tune_record: one level shape the autotuner has decided."""
    _fields_ = [("kind", ctypes.c_int32), ("fp64", ctypes.c_int32), ("half", ctypes.c_int32),
                ("rows_log2", ctypes.c_int32), ("per_row", ctypes.c_int32), ("ncand", ctypes.c_int32),
                ("choice", ctypes.c_int32), ("samples", ctypes.c_int32), ("cand", ctypes.c_int32 * 4),
                ("us", ctypes.c_double * 4), ("graphs_log2", ctypes.c_int32), ("pad_", ctypes.c_int32)]
    KINDS = ("smoothing", "prolongation", "restriction", "dense")

    def as_dict(self):
        n = self.ncand
        return {"kind": self.KINDS[self.kind], "fp64": bool(self.fp64), "half": bool(self.half),
                "rows_log2": self.rows_log2, "per_row": self.per_row, "graphs_log2": self.graphs_log2,
                "candidates": list(self.cand[:n]),
                "us": [round(x, 2) for x in self.us[:n]], "choice": self.cand[self.choice]}


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("steps", ctypes.c_int64), ("looks", ctypes.c_int64), ("assembled", ctypes.c_int64),
                ("max_active", ctypes.c_int64), ("step_ms", ctypes.c_double), ("rebuild_ms", ctypes.c_double),
                ("setup_ms", ctypes.c_double), ("assemble_ms", ctypes.c_double), ("setup_groups", ctypes.c_int64),
                ("rebuilds", ctypes.c_int64), ("restarts", ctypes.c_int64), ("solves_abandoned", ctypes.c_int64),
                ("breakdowns", ctypes.c_int64), ("diverged", ctypes.c_int64), ("judged", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# ---- the judge and break rounds (World::Impl::judge, world.cpp:1162-1277) ------------------------
SVXG_JUDGE_STRUCTURE, SVXG_JUDGE_PIECE = 0, 1
(SVXG_CASCADE_SETTLED, SVXG_CASCADE_ROUNDS, SVXG_CASCADE_UNSOLVED, SVXG_CASCADE_GONE, SVXG_CASCADE_RESEED,
 SVXG_CASCADE_LIMIT) = range(6)
CASCADE_ENDS = ("settled", "rounds", "unsolved", "gone", "reseed", "limit")

HINGE_CUT = np.dtype([
    ("a", "<i4"), ("b", "<i4"), ("p", "<f8", (3,)), ("axis", "<f8", (3,)), ("n", "<f8", (3,)),
    ("mp", "<f8"), ("pull", "<f8"),
], align=True)
BREAK = np.dtype([
    ("bond", "<i4"), ("round", "<i4"), ("phi", "<f8"), ("mode", "u1"), ("hinge", "u1"), ("pad_", "u1", (6,)),
    ("cut", HINGE_CUT),
], align=True)
assert HINGE_CUT.itemsize == 96 and BREAK.itemsize == 120


class JudgeOptions(ctypes.Structure):
    """WorldConfig's judge settings (core/include/svx/world/world.hpp:114-183)."""
    _fields_ = [("fragility", ctypes.c_double), ("dif", ctypes.c_double), ("break_band", ctypes.c_double),
                ("quota_fraction", ctypes.c_double), ("hinge_shape", ctypes.c_double),
                ("ductile", ctypes.c_uint64 * 2), ("rule", ctypes.c_int32), ("max_breaks", ctypes.c_int32),
                ("max_rounds", ctypes.c_int32), ("plastic_hinges", ctypes.c_int32),
                ("restart_iters", ctypes.c_int32), ("rebuild_iters", ctypes.c_int32)]


class JudgeState(ctypes.Structure):
    """A structure's cascade state (Structure::shock, transient, rounds)."""
    _fields_ = [("shock", ctypes.c_int32), ("transient", ctypes.c_int32), ("rounds", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]

    def as_dict(self):
        return {"shock": bool(self.shock), "transient": bool(self.transient), "rounds": self.rounds}


class JudgeResult(ctypes.Structure):
    _fields_ = [("max_phi", ctypes.c_double), ("over", ctypes.c_int32), ("picked", ctypes.c_int32),
                ("hinges", ctypes.c_int32), ("pad_", ctypes.c_int32)]

    def as_dict(self):
        return {"max_phi": self.max_phi, "over": self.over, "picked": self.picked, "hinges": self.hinges}


class ServerOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_active", ctypes.c_int32), ("rtol", ctypes.c_double),
                ("maxit", ctypes.c_int32), ("amg_min_nodes", ctypes.c_int32), ("window", ctypes.c_int32),
                ("restart_iters", ctypes.c_int32), ("rebuild_iters", ctypes.c_int32), ("setup_workers", ctypes.c_int32),
                ("solve_restarts", ctypes.c_int32), ("restart_diverging", ctypes.c_int32), ("judge", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("judge_opt", JudgeOptions)]


class ServeDone(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("problem", ctypes.c_void_p), ("result", PcgResult),
                ("status", ctypes.c_int32), ("assembled", ctypes.c_int32), ("t_submit_ms", ctypes.c_double),
                ("t_ready_ms", ctypes.c_double), ("t_admit_ms", ctypes.c_double), ("t_done_ms", ctypes.c_double),
                ("restarts", ctypes.c_int32), ("abandoned", ctypes.c_int32), ("breakdowns", ctypes.c_int32),
                ("detached", ctypes.c_int32), ("judge", JudgeResult)]


class CascadeResult(ctypes.Structure):
    _fields_ = [("last", PcgResult), ("max_phi", ctypes.c_double), ("end", ctypes.c_int32),
                ("rounds", ctypes.c_int32), ("breaks", ctypes.c_int32), ("retired", ctypes.c_int32),
                ("solves", ctypes.c_int32), ("iters", ctypes.c_int32), ("rebuilds", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]

    def as_dict(self):
        return {"last": self.last.as_dict(), "max_phi": self.max_phi, "end": CASCADE_ENDS[self.end],
                "rounds": self.rounds, "breaks": self.breaks, "retired": self.retired, "solves": self.solves,
                "iters": self.iters, "rebuilds": self.rebuilds}


assert ctypes.sizeof(JudgeOptions) == 80 and ctypes.sizeof(JudgeResult) == 24
assert ctypes.sizeof(CascadeResult) == 64 and ctypes.sizeof(JudgeState) == 16


def ductile_mask(ids):
    """svxg_judge_options.ductile from a list of material ids (Material::ductile, material.hpp:66)."""
    m = [0, 0]
    for i in ids:
        if not 0 <= int(i) < 128:
            raise ValueError(f"material id {i} out of range")
        m[int(i) >> 6] |= 1 << (int(i) & 63)
    return m


# ---- bond sections from voxel faces (world_internal.hpp:105-125, 129-249) ------------------------
SECTION_HEAD = np.dtype([
    ("a", "<i4"), ("b", "<i4"), ("axis", "<i4"), ("sign", "<i4"), ("face0", "<i4"), ("nfaces", "<i4"),
    ("ma", "u1"), ("mb", "u1"), ("pad_", "u1", (6,)), ("strength_a", "<f8"), ("strength_b", "<f8"),
], align=True)
SECTION_FACE = np.dtype([("p", "<i4", (3,)), ("axis", "i1"), ("sign", "i1"), ("pad_", "u1", (2,))], align=True)
assert SECTION_HEAD.itemsize == 48 and SECTION_FACE.itemsize == 16


# ---- the general case: placed lattices and junction samples (world_internal.hpp:78-125, 250-441) ----
SECTION_HEAD_GENERAL = np.dtype([
    ("a", "<i4"), ("b", "<i4"), ("axis", "<i4"), ("sign", "<i4"), ("face0", "<i4"), ("nfaces", "<i4"),
    ("sample0", "<i4"), ("nsamples", "<i4"), ("grid", "<u2"), ("ma", "u1"), ("mb", "u1"), ("pad_", "u1", (4,)),
    ("strength_a", "<f8"), ("strength_b", "<f8"),
], align=True)
JUNCTION_SAMPLE = np.dtype([
    ("v", "<i4", (3,)), ("o", "<i4", (3,)), ("vg", "<u2"), ("og", "<u2"), ("face", "u1"), ("sub", "u1"),
    ("kind", "u1"), ("sign", "i1"), ("weight", "<f4"),
], align=True)
assert SECTION_HEAD_GENERAL.itemsize == 56 and JUNCTION_SAMPLE.itemsize == 36


class Lattice(ctypes.Structure):
    """svxg_lattice: a lattice placed in the bond's frame (to(s) = off + R s; identity: unplaced) with
    its voxels (Vox codes, x fastest)."""
    _fields_ = [("off", ctypes.c_double * 3), ("R", ctypes.c_double * 9), ("h", ctypes.c_double),
                ("identity", ctypes.c_int32), ("lo", ctypes.c_int32 * 3), ("dim", ctypes.c_int32 * 3),
                ("pad_", ctypes.c_int32), ("vox", ctypes.c_void_p), ("damage", ctypes.c_void_p)]


class VoxelBox(ctypes.Structure):
    """svxg_voxel_box: Vox codes (0 air, material id + 1, 0x80 anchored), x fastest."""
    _fields_ = [("lo", ctypes.c_int32 * 3), ("dim", ctypes.c_int32 * 3), ("h", ctypes.c_double),
                ("vox", ctypes.c_void_p), ("damage", ctypes.c_void_p)]


# ---- a structure's extraction (World::Impl::extract + append_nodes, world.cpp:425-865) ----------
FRAGMENT = np.dtype([
    ("com", "<f8", (3,)), ("mass", "<f8"), ("strength", "<f8"), ("owner", "<i8"), ("first", "<i4", (3,)),
    ("mat", "u1"), ("pad_", "u1", (3,)),
], align=True)
assert FRAGMENT.itemsize == 64


class ExtractOptions(ctypes.Structure):
    """svxg_extract_options (WorldConfig's extraction settings, world.hpp:142-148, 234-265)."""
    _fields_ = [("max_nodes", ctypes.c_int32), ("cluster_nodes", ctypes.c_int32), ("max_radius", ctypes.c_double),
                ("fresh_from", ctypes.c_int64), ("cluster_cubes", ctypes.c_int32), ("outside", ctypes.c_int32)]


class ExtractResult(ctypes.Structure):
    """svxg_extract_result: in, the room of the outputs; out, the counts and the walk's flags."""
    _fields_ = [("members", ctypes.c_int32), ("nodes", ctypes.c_int32), ("bonds", ctypes.c_int32),
                ("faces", ctypes.c_int32), ("cell", ctypes.c_int32), ("truncated", ctypes.c_int32),
                ("any_support", ctypes.c_int32), ("touched_unloaded", ctypes.c_int32)]


assert ctypes.sizeof(ExtractOptions) == 32 and ctypes.sizeof(ExtractResult) == 32


def extract_options(**kw):
    """svxg_extract_options at the reference's defaults (max_nodes 60000, cluster_nodes 2500,
    max_radius 60 m, no fresh owners, cluster_cubes off, air outside the box), then kw."""
    o = ExtractOptions(max_nodes=60000, cluster_nodes=2500, max_radius=60.0, fresh_from=(1 << 63) - 1,
                       cluster_cubes=0, outside=0)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def ptr(a):
    """The data pointer of a numpy array (None for None)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def reference_options(**kw):
    """StressOptions{} with AmgOptions{} as the reference defaults them."""
    o = StressOptions()
    o.rtol = 2e-3
    o.amg_min_nodes = 160
    o.precision = SVXG_FP64
    o.amg.strength = 0.1
    o.amg.coarse_scale = 0.6
    o.amg.coarse_max = 48
    o.amg.smoothed = 1
    o.amg.max_levels = 16
    o.amg.sweeps = 1
    o.amg.cheb_degree = 2
    o.amg.coarsen_dense = 1
    o.amg.reaggregate = 1
    for k, v in kw.items():
        if hasattr(o.amg, k) and k not in ("rtol",):
            setattr(o.amg, k, v)
        else:
            setattr(o, k, v)
    return o
