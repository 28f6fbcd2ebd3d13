"""Python mirror of the reference's StressProblem (core/include/svx/stress/stress.hpp:74-160) over
the device library's C ABI (include/svxg.h, libsvxg.so).

Same names, argument meaning and results as the reference: assemble(options) -> bool,
solve(f, u, rtol, maxit, warm) -> (u, PcgResult), begin / iterate / current, bond_load,
remove_bond, retire_nodes, plus the batched entry solve_batch(). Nothing here computes: every call
goes to the device; a missing or unloadable libsvxg.so raises.
"""
import ctypes
import os

import numpy as np

from .abi import (BOND, BOND_LOAD, NODE, SVXG_FP64, SVXG_MIXED, SVXG_OK, SVXG_SINGULAR, PcgResult, ProblemInfo,
                  StressOptions, reference_options,
                  BREAK, CascadeResult, JudgeOptions, JudgeResult, JudgeState, TuneRecord, ductile_mask)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsvxg.so")
_LIB = None


class SvxgError(RuntimeError):
    pass


def lib():
    """The loaded device library (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SvxgError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        # A server runs its engine and its setup workers on streams of their own, beside the
        # context's; with HIP's default of 4 hardware queues a process's streams beyond 4 share
        # queues, and setup work then waits behind solves (a 60 Hz stream's p50 tick 13.3-13.8 ->
        # 11.5 ms with 8). The HIP runtime reads this when it starts: before the first HIP call.
        try:
            queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        except ValueError:
            queues = 4
        if queues < 8:
            os.environ["GPU_MAX_HW_QUEUES"] = "8"
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        sig = {
            "svxg_ctx_create": ([ctypes.c_int, i64, ctypes.POINTER(vp)], ctypes.c_int),
            "svxg_ctx_destroy": ([vp], ctypes.c_int),
            "svxg_ctx_bytes_in_use": ([vp], i64),
            "svxg_last_error": ([], ctypes.c_char_p),
            "svxg_default_options": ([vp], None),
            "svxg_problem_create": ([vp, vp, i32, vp, i32, vp, i32, ctypes.POINTER(vp)], ctypes.c_int),
            "svxg_problem_destroy": ([vp], ctypes.c_int),
            "svxg_problem_assemble": ([vp, vp], ctypes.c_int),
            "svxg_problem_release": ([vp], ctypes.c_int),
            "svxg_problem_invalidate": ([vp], ctypes.c_int),
            "svxg_problem_solve": ([vp, vp, vp, f64, i32, i32, vp], ctypes.c_int),
            "svxg_problem_begin": ([vp, vp, vp], ctypes.c_int),
            "svxg_problem_iterate": ([vp, i32, f64, vp], ctypes.c_int),
            "svxg_problem_current": ([vp, vp], ctypes.c_int),
            "svxg_problem_remove_bonds": ([vp, vp, i32], ctypes.c_int),
            "svxg_problem_retire_nodes": ([vp, vp, i32], ctypes.c_int),
            "svxg_problem_set_bond_broken": ([vp, i32, i32], ctypes.c_int),
            "svxg_problem_bond_loads": ([vp, vp, vp], ctypes.c_int),
            "svxg_problem_bond_checks": ([vp, vp, f64, vp, vp], ctypes.c_int),
            "svxg_problem_bond_checks_noalloc": ([vp, vp, f64, vp, vp], ctypes.c_int),
            "svxg_add_force": ([vp, i32, vp, vp, vp], None),
            "svxg_extend_warm_start": ([vp, i32, vp, i32, vp, i32], None),
            "svxg_problem_info_get": ([vp, vp], ctypes.c_int),
            "svxg_problem_get_bonds": ([vp, vp], ctypes.c_int),
            "svxg_problem_get_nodes": ([vp, vp], ctypes.c_int),
            "svxg_solve_batch": ([vp, vp, i32, vp, vp, f64, i32, i32, vp], ctypes.c_int),
            "svxg_problem_get_matrix": ([vp, vp, vp, vp], ctypes.c_int),
            "svxg_problem_get_diag_inverse": ([vp, ctypes.c_int32, vp, vp], ctypes.c_int),
            "svxg_problem_apply_matrix": ([vp, vp, vp, i32], ctypes.c_int),
            "svxg_problem_apply_preconditioner": ([vp, vp, vp], ctypes.c_int),
            "svxg_problem_last_timing": ([vp, ctypes.POINTER(f64), ctypes.POINTER(f64)], ctypes.c_int),
            "svxg_problem_set_options": ([vp, vp], ctypes.c_int),
            "svxg_assemble_batch": ([vp, vp, i32, vp], ctypes.c_int),
            "svxg_ctx_set_profiling": ([vp, i32], ctypes.c_int),
            "svxg_ctx_set_resident_rows": ([vp, i32], ctypes.c_int),
            "svxg_ctx_set_step_graphs": ([vp, i32], ctypes.c_int),
            "svxg_ctx_set_autotune": ([vp, i32], ctypes.c_int),
            "svxg_ctx_tuned": ([vp, vp, i32, ctypes.POINTER(i32)], ctypes.c_int),
            "svxg_ctx_kernel_stats": ([vp, i32, ctypes.POINTER(i64), ctypes.POINTER(f64), ctypes.POINTER(f64)],
                                      ctypes.c_int),
            "svxg_server_default_options": ([vp], None),
            "svxg_server_create": ([vp, vp, ctypes.POINTER(vp)], ctypes.c_int),
            "svxg_server_destroy": ([vp], ctypes.c_int),
            "svxg_server_submit": ([vp, vp, vp, vp, i32, vp, i32, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "svxg_server_poll": ([vp, vp, i32, ctypes.POINTER(i32), i32], ctypes.c_int),
            "svxg_server_submit_many": ([vp, i32, vp, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "svxg_server_release": ([vp, vp], ctypes.c_int),
            "svxg_server_now_ms": ([vp], f64),
            "svxg_server_stats_get": ([vp, vp], ctypes.c_int),
            "svxg_server_picks": ([vp, vp, vp, i32, ctypes.POINTER(i32)], ctypes.c_int),
            "svxg_default_judge_options": ([vp], None),
            "svxg_problem_set_judge_baseline": ([vp, vp, vp], ctypes.c_int),
            "svxg_problem_get_judge_baseline": ([vp, vp, vp], ctypes.c_int),
            "svxg_judge_batch": ([vp, vp, i32, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "svxg_break_rounds_batch": ([vp, vp, i32, vp, vp, f64, i32, vp, i32, vp, vp, vp, vp, vp], ctypes.c_int),
            "svxg_bond_sections": ([vp, vp, i32, vp, i32, vp, i32, vp, vp, i32, vp], ctypes.c_int),
            "svxg_bond_sections_general": ([vp, vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, i32, vp, i32, vp],
                                           ctypes.c_int),
            "svxg_lattice_graph": ([vp, vp, vp, i32, vp, i32, vp, ctypes.POINTER(i32), ctypes.POINTER(i32), vp, vp, vp],
                                   ctypes.c_int),
            "svxg_lattice_graph_clustered": ([vp, vp, vp, i32, i32, i32, vp, i32, vp, ctypes.POINTER(i32), vp,
                                              ctypes.POINTER(i32), ctypes.POINTER(i32), vp, vp, vp], ctypes.c_int),
            "svxg_extract_default_options": ([vp], None),
            "svxg_extract_structure": ([vp, vp, vp, vp, vp, i32, i32, vp, vp, i32, vp] + [vp] * 8, ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _check(status, what):
    if status != SVXG_OK:
        msg = lib().svxg_last_error().decode(errors="replace")
        raise SvxgError(f"{what}: status {status}: {msg}")


def default_options(**kw):
    """StressOptions{} (stress.hpp:68-72) as the device library defaults them."""
    o = StressOptions()
    lib().svxg_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if k in ("rtol", "amg_min_nodes", "precision", "mixed_fp64_rows", "vcycle_fp16_rows"):
            setattr(o, k, v)
        else:
            setattr(o.amg, k, v)
    return o


class Context:
    """One GPU and its HBM budget (bytes; 0 = 200 GB)."""

    def __init__(self, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        _check(lib().svxg_ctx_create(device, hbm_budget, ctypes.byref(h)), "svxg_ctx_create")
        self.h = h

    def bytes_in_use(self):
        return lib().svxg_ctx_bytes_in_use(self.h)

    def set_resident_rows(self, rows):
        """Graphs up to `rows` free nodes step in one workgroup each (0: none); results are unchanged."""
        _check(lib().svxg_ctx_set_resident_rows(self.h, int(rows)), "set_resident_rows")

    def set_step_graphs(self, mode):
        """Tiled solve steps recorded as a HIP graph and replayed: 0 never, 1 large batches
        (default), 2 always. Results are unchanged."""
        _check(lib().svxg_ctx_set_step_graphs(self.h, int(mode)), "set_step_graphs")

    def set_profiling(self, on):
        _check(lib().svxg_ctx_set_profiling(self.h, int(bool(on))), "set_profiling")

    def set_autotune(self, on):
        """Tile shapes tuned online per level shape (default on) or the untuned defaults; results are
        unchanged either way."""
        _check(lib().svxg_ctx_set_autotune(self.h, int(bool(on))), "set_autotune")

    def tuned(self):
        """The level shapes the autotuner has decided: a list of dicts (kind, shape, candidates, their
        mean launch times in microseconds, the choice)."""
        n = ctypes.c_int32()
        _check(lib().svxg_ctx_tuned(self.h, None, 0, ctypes.byref(n)), "tuned")
        buf = (TuneRecord * max(1, n.value))()
        _check(lib().svxg_ctx_tuned(self.h, buf, n.value, ctypes.byref(n)), "tuned")
        return [buf[k].as_dict() for k in range(n.value)]

    def kernel_stats(self, kind):
        """(launches, device ms, algorithmic bytes) of a profiled kernel kind since the last call."""
        n, ms, b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _check(lib().svxg_ctx_kernel_stats(self.h, int(kind), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
               "kernel_stats")
        return n.value, ms.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().svxg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT_CTX = None


def default_context():
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class StressProblem:
    """A fragment graph's stress problem on the device (StressProblem, stress.hpp:74-160)."""

    def __init__(self, nodes, bonds, mats, ctx=None, options=None):
        self.ctx = ctx or default_context()
        self._nodes = np.ascontiguousarray(nodes, NODE)
        self._bonds = np.ascontiguousarray(bonds, BOND)
        self.mats = np.ascontiguousarray(mats)
        h = ctypes.c_void_p()
        _check(lib().svxg_problem_create(self.ctx.h, self._nodes.ctypes.data, len(self._nodes), self._bonds.ctypes.data,
                                         len(self._bonds), self.mats.ctypes.data, len(self.mats), ctypes.byref(h)),
               "svxg_problem_create")
        self.h = h
        self.options = default_options()
        if options is not None:
            self.set_options(options)

    def set_options(self, options):
        """The options of the next assemble (batched assembly uses them)."""
        self.options = options
        _check(lib().svxg_problem_set_options(self.h, ctypes.byref(options)), "set_options")

    def close(self):
        if getattr(self, "h", None):
            lib().svxg_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return len(self._nodes)

    @property
    def nodes(self):
        out = self._nodes.copy()
        _check(lib().svxg_problem_get_nodes(self.h, out.ctypes.data), "get_nodes")
        return out

    @property
    def bonds(self):
        out = self._bonds.copy()
        _check(lib().svxg_problem_get_bonds(self.h, out.ctypes.data), "get_bonds")
        return out

    def assemble(self, options=None):
        """StressProblem::assemble (stress.cpp:134-176): False where the reference returns false."""
        if options is not None:
            self.options = options
        st = lib().svxg_problem_assemble(self.h, ctypes.byref(self.options))
        if st == SVXG_SINGULAR:
            return False
        _check(st, "svxg_problem_assemble")
        return True

    def release(self):
        _check(lib().svxg_problem_release(self.h), "release")

    def invalidate(self):
        _check(lib().svxg_problem_invalidate(self.h), "invalidate")

    def solve(self, f, u=None, rtol=1e-8, maxit=-1, warm=False):
        """StressProblem::solve (stress.cpp:348-360): (u over all nodes, PcgResult as a dict)."""
        f = np.ascontiguousarray(f, np.float64)
        u = np.zeros(6 * self.n) if u is None else np.array(u, np.float64)
        r = PcgResult()
        _check(lib().svxg_problem_solve(self.h, f.ctypes.data, u.ctypes.data, rtol, int(maxit), int(bool(warm)),
                                        ctypes.byref(r)), "svxg_problem_solve")
        return u, r.as_dict()

    def begin(self, f, u0=None):
        f = np.ascontiguousarray(f, np.float64)
        u0 = None if u0 is None else np.ascontiguousarray(u0, np.float64)
        _check(lib().svxg_problem_begin(self.h, f.ctypes.data, None if u0 is None else u0.ctypes.data), "begin")

    def iterate(self, maxit, rtol):
        r = PcgResult()
        _check(lib().svxg_problem_iterate(self.h, int(maxit), float(rtol), ctypes.byref(r)), "iterate")
        return r.as_dict()

    def current(self):
        u = np.zeros(6 * self.n)
        _check(lib().svxg_problem_current(self.h, u.ctypes.data), "current")
        return u

    def remove_bond(self, b):
        self.remove_bonds([b])

    def remove_bonds(self, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        _check(lib().svxg_problem_remove_bonds(self.h, ids.ctypes.data, len(ids)), "remove_bonds")

    def retire_nodes(self, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        _check(lib().svxg_problem_retire_nodes(self.h, ids.ctypes.data, len(ids)), "retire_nodes")

    def set_bond_broken(self, b, broken):
        _check(lib().svxg_problem_set_bond_broken(self.h, int(b), int(bool(broken))), "set_bond_broken")

    def bond_loads(self, u=None):
        """bond_load of every bond (stress.cpp:477-502); u None: the last solve's, resident."""
        out = np.zeros(len(self._bonds), BOND_LOAD)
        uu = None if u is None else np.ascontiguousarray(u, np.float64)
        _check(lib().svxg_problem_bond_loads(self.h, None if uu is None else uu.ctypes.data, out.ctypes.data),
               "bond_loads")
        return out

    def bond_load(self, b, u=None):
        return self.bond_loads(u)[b]

    def bond_checks(self, u=None, fragility=1.0, out=None):
        """(phi, mode) of every bond: bond_utilization after bond_load (stress.cpp:14-57). out: a
        (phi float64, mode uint8) pair of contiguous arrays of the bonds' length to write into (a
        caller checking every tick keeps its arrays)."""
        if out is None:
            phi = np.zeros(len(self._bonds))
            mode = np.zeros(len(self._bonds), np.uint8)
        else:
            phi, mode = out
            if (phi.dtype != np.float64 or mode.dtype != np.uint8 or phi.shape != (len(self._bonds),)
                    or mode.shape != (len(self._bonds),) or not phi.flags.c_contiguous or not mode.flags.c_contiguous):
                raise ValueError("bond_checks: out must be contiguous float64 and uint8 arrays of the bonds' length")
        uu = None if u is None else np.ascontiguousarray(u, np.float64)
        _check(lib().svxg_problem_bond_checks(self.h, None if uu is None else uu.ctypes.data, float(fragility),
                                              phi.ctypes.data, mode.ctypes.data), "bond_checks")
        return phi, mode

    def info(self):
        i = ProblemInfo()
        _check(lib().svxg_problem_info_get(self.h, ctypes.byref(i)), "info")
        return i.as_dict()

    def work_per_iteration(self):
        return self.info()["work_per_iteration"]

    def matrix(self):
        inf = self.info()
        rowptr = np.zeros(inf["nfree"] + 1, np.int32)
        col = np.zeros(inf["blocks"], np.int32)
        val = np.zeros((inf["blocks"], 6, 6))
        _check(lib().svxg_problem_get_matrix(self.h, rowptr.ctypes.data, col.ctypes.data, val.ctypes.data), "matrix")
        return rowptr, col, val

    def diag_inverse(self, level=0):
        """(blocks [rows, 6, 6], held in FP64) of a level's inverse diagonal (amg.cpp:227-244)."""
        inf = self.info()
        rows = inf["nfree"] if level == 0 else inf["level_rows"][level]  # (block-Jacobi: no hierarchy)
        out = np.zeros((rows, 6, 6))
        f64 = ctypes.c_int32(0)
        _check(lib().svxg_problem_get_diag_inverse(self.h, int(level), out.ctypes.data, ctypes.byref(f64)),
               "diag_inverse")
        return out, bool(f64.value)

    def apply_matrix(self, x, fp32=False):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        _check(lib().svxg_problem_apply_matrix(self.h, x.ctypes.data, y.ctypes.data, int(bool(fp32))), "apply_matrix")
        return y

    def apply_preconditioner(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        _check(lib().svxg_problem_apply_preconditioner(self.h, r.ctypes.data, z.ctypes.data), "apply_preconditioner")
        return z

    def set_judge_baseline(self, loads=None, has=None):
        """The judged loads a shock is measured from (World::Impl::judged_, world.cpp:1176-1190):
        loads (BOND_LOAD per bond) and has (per bond, 0: none); None clears them."""
        if loads is None:
            _check(lib().svxg_problem_set_judge_baseline(self.h, None, None), "set_judge_baseline")
            return
        loads = np.ascontiguousarray(loads, BOND_LOAD)
        has = np.ascontiguousarray(has, np.uint8)
        _check(lib().svxg_problem_set_judge_baseline(self.h, loads.ctypes.data, has.ctypes.data), "set_judge_baseline")

    def judge_baseline(self):
        m = len(self._bonds)
        loads = np.zeros(m, BOND_LOAD)
        has = np.zeros(m, np.uint8)
        _check(lib().svxg_problem_get_judge_baseline(self.h, loads.ctypes.data, has.ctypes.data), "judge_baseline")
        return loads, has

    def last_timing(self):
        s, a = ctypes.c_double(), ctypes.c_double()
        _check(lib().svxg_problem_last_timing(self.h, ctypes.byref(s), ctypes.byref(a)), "last_timing")
        return {"solve_ms": s.value, "setup_ms": a.value}


def add_force(f, node, c, F, at):
    """StressProblem::add_force (stress.cpp:504-513) on a host load vector."""
    c, F, at = (np.ascontiguousarray(v, np.float64) for v in (c, F, at))
    lib().svxg_add_force(f.ctypes.data, int(node), c.ctypes.data, F.ctypes.data, at.ctypes.data)


def assemble_batch(problems, ctx=None):
    """Assemble many problems in one set of device launches; returns each one's success (bool)."""
    ctx = ctx or problems[0].ctx
    n = len(problems)
    ph = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    st = (ctypes.c_int32 * n)()
    _check(lib().svxg_assemble_batch(ctx.h, ph, n, st), "svxg_assemble_batch")
    return [s == SVXG_OK for s in st]


def extend_warm_start(nodes, bonds, u, first_new):
    """StressProblem::extend_warm_start (stress.hpp:115): u (6 per node) of the nodes from first_new
    on, carried rigidly from their bonded older neighbours (a copy is returned)."""
    nodes = np.ascontiguousarray(nodes, NODE)
    bonds = np.ascontiguousarray(bonds, BOND)
    out = np.array(u, np.float64)
    lib().svxg_extend_warm_start(nodes.ctypes.data, len(nodes), bonds.ctypes.data, len(bonds), out.ctypes.data,
                                 int(first_new))
    return out


def solve_batch(problems, fs, us=None, rtol=1e-8, maxit=-1, warm=False, ctx=None):
    """Solve many problems in one set of device launches; returns (list of u, list of results)."""
    ctx = ctx or problems[0].ctx
    n = len(problems)
    fs = [np.ascontiguousarray(f, np.float64) for f in fs]
    if us is None:
        us = [np.zeros(6 * p.n) for p in problems]
    else:
        us = [np.array(u, np.float64) for u in us]
    ph = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    fp = (ctypes.c_void_p * n)(*[f.ctypes.data for f in fs])
    up = (ctypes.c_void_p * n)(*[u.ctypes.data for u in us])
    res = (PcgResult * n)()
    _check(lib().svxg_solve_batch(ctx.h, ph, n, fp, up, rtol, int(maxit), int(bool(warm)), res), "svxg_solve_batch")
    return us, [r.as_dict() for r in res]


def default_judge_options(**kw):
    """svxg_judge_options as the library defaults them (WorldConfig, world.hpp:114-183); ductile
    takes a list of material ids."""
    o = JudgeOptions()
    lib().svxg_default_judge_options(ctypes.byref(o))
    for k, v in kw.items():
        if k == "ductile":
            o.ductile[0], o.ductile[1] = ductile_mask(v)
        else:
            setattr(o, k, v)
    return o


def _states(states, n):
    st = (JudgeState * n)()
    for k in range(n):
        d = (states[k] if states is not None else None) or {}
        st[k].shock = int(bool(d.get("shock", False)))
        st[k].transient = int(bool(d.get("transient", False)))
        st[k].rounds = int(d.get("rounds", 0))
    return st


def judge_batch(problems, states=None, opts=None, us=None, want_phi=False, ctx=None):
    """World::Impl::judge without side effects for many problems (svxg_judge_batch): per problem
    a dict of max_phi, over, picked, hinges, picks (BREAK records in order) and phi (if asked)."""
    ctx = ctx or problems[0].ctx
    n = len(problems)
    opts = opts or default_judge_options()
    ph = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    ua = None
    if us is not None:
        us = [None if u is None else np.ascontiguousarray(u, np.float64) for u in us]
        ua = (ctypes.c_void_p * n)(*[None if u is None else u.ctypes.data for u in us])
    st = _states(states, n)
    res = (JudgeResult * n)()
    picks = [np.zeros(opts.max_breaks, BREAK) for _ in problems]
    pa = (ctypes.c_void_p * n)(*[p.ctypes.data for p in picks])
    phis = [np.zeros(len(p._bonds)) for p in problems] if want_phi else None
    phia = (ctypes.c_void_p * n)(*[f.ctypes.data for f in phis]) if want_phi else None
    _check(lib().svxg_judge_batch(ctx.h, ph, n, ua, ctypes.byref(opts), st, res, pa, phia), "svxg_judge_batch")
    out = []
    for k in range(n):
        d = res[k].as_dict()
        d["picks"] = picks[k][:d["picked"]].copy()
        if want_phi:
            d["phi"] = phis[k]
        out.append(d)
    return out


def break_rounds_batch(problems, fs, us=None, opts=None, states=None, rtol=1e-8, maxit=5000, max_call_rounds=0,
                       capacity=None, ctx=None):
    """The break-round cascade of many structures (svxg_break_rounds_batch): returns (us, results),
    each result a dict with the cascade's counters, 'breaks' (BREAK records), 'retired' (node ids)
    and 'state' (the advanced JudgeState)."""
    ctx = ctx or problems[0].ctx
    n = len(problems)
    opts = opts or default_judge_options()
    fs = [np.ascontiguousarray(f, np.float64) for f in fs]
    us = [np.zeros(6 * p.n) if us is None or us[k] is None else np.array(us[k], np.float64)
          for k, p in enumerate(problems)]
    ph = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    fp = (ctypes.c_void_p * n)(*[f.ctypes.data for f in fs])
    up = (ctypes.c_void_p * n)(*[u.ctypes.data for u in us])
    st = _states(states, n)
    res = (CascadeResult * n)()
    cap = [int(capacity) if capacity is not None else len(p._bonds) for p in problems]
    brs = [np.zeros(max(1, c), BREAK) for c in cap]
    ba = (ctypes.c_void_p * n)(*[b.ctypes.data for b in brs])
    ca = (ctypes.c_int32 * n)(*cap)
    ret = [np.zeros(max(1, p.n), np.int32) for p in problems]
    ra = (ctypes.c_void_p * n)(*[r.ctypes.data for r in ret])
    _check(lib().svxg_break_rounds_batch(ctx.h, ph, n, fp, up, rtol, int(maxit), ctypes.byref(opts),
                                         int(max_call_rounds), st, res, ba, ca, ra), "svxg_break_rounds_batch")
    out = []
    for k in range(n):
        d = res[k].as_dict()
        d["breaks"] = brs[k][:min(d["breaks"], cap[k])].copy()
        d["retired"] = ret[k][:min(d["retired"], problems[k].n)].copy()
        d["state"] = st[k].as_dict()
        out.append(d)
    return us, out


__all__ = ["StressProblem", "Context", "default_context", "default_options", "reference_options", "solve_batch",
           "assemble_batch",
           "add_force", "lib", "SvxgError", "SVXG_FP64", "SVXG_MIXED", "default_judge_options", "judge_batch",
           "break_rounds_batch"]
