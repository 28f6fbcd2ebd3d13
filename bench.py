"""Benchmark of the stress solve on one MI355X.

Workload (the metric's single-GPU configuration): one irregular block of 64 x 64 x 48 fragments with
15 % voids (about 167k nodes, 1.0M DOFs, stiffness ratio up to 20:1), solved to a relative residual
of 1e-8 and bond-checked. A step is one solve from a cold start plus the bond checks on the
device-resident solution. The headline runs the mixed path (FP64 PCG and K products around the FP32
V-cycle); the plain FP64 path (FP64 V-cycle) and the mixed split with FP32 K products inside FP64
iterative refinement are measured beside it. Further legs, each reported in the same line: the
batched uneven graphs (4096 pieces), the serving stream (60 Hz ticks), the city chunk (400
resident buildings with ground-storey collapses), the block's break-round cascade (strengths
scaled so that its largest utilisation is 3, solved, judged and broken until it settles) and a
structure's graph built from voxels (a 128 x 128 x 96 Voronoi block: bond sections on the device).

value: solved DOFs / (device time of the solve + wall time of the checks), inputs resident and the
checks' phi and modes written to device arrays; the wall rate with the load, the solution and phi
crossing PCIe is reported beside it (wall_ms_pcie_inclusive), never as value.
roofline: the dominant kernel's algorithmic bytes per launch / its average launch time, measured
live with HIP events on the library's stream during as many solves again after the timed steps.
cpu_baseline: the reference's own solver (compiled from its sources into oracle/_ref) on the same
workload: the median of --cpu-runs runs of assemble + cold solve on every host thread of this
process's share (setup and solve timed apart), with a one-thread figure beside it; cpu_baselines
holds the same for the tower (C1) and the batched leg's plain sample (the piece path).

  python bench.py [--steps K] [--warmup W] [--precision mixed|fp64] [--no-cpu] [--gpus 1]
  python bench.py --only batch|serve|city|cascade|graph ...   (one leg alone, for profiling)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes of one live launch of a kernel on a workload, from two rocprofv3 counter passes of this
# script (FETCH_SIZE scaled by 2.0, the factor a plain 1 GiB copy reads it with; WRITE_SIZE as read)
MEASURED_TRAFFIC = {("k_cheb_step<float, 1, false> (level 0, FP16 block copies)", 64, 64, 48): 114.2e6,
                    ("k_cheb_step<float, 1, false> (level 0)", 64, 64, 48): 202.7e6}
METRIC = "PCG-solved DOFs/s at 1e-8 residual, batched vs plain; p50/p99 tick latency"
PRECISION_MIXED = ("FP64 PCG (FP64 K products, dots and vector updates) around an FP32-arithmetic V-cycle with an "
                   "FP64 dense coarsest solve")
PRECISION_FP64 = "FP64 PCG around an FP64 V-cycle (every product, smoother and vector in FP64)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


AMG_KW = {}  # (device multigrid options of every leg, from the command line)


class DeviceArray:
    """A device buffer from the HIP runtime the library itself loads (hipMalloc through ctypes): the
    bond checks' outputs kept in HBM, as a caller whose judge runs on the device keeps them."""
    _hip = None

    def __init__(self, nbytes):
        import ctypes
        if DeviceArray._hip is None:
            L = ctypes.CDLL("libamdhip64.so.7")
            L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            L.hipFree.argtypes = [ctypes.c_void_p]
            L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            DeviceArray._hip = L
        self.p, self.nbytes = ctypes.c_void_p(), max(1, int(nbytes))
        if DeviceArray._hip.hipMalloc(ctypes.byref(self.p), self.nbytes) != 0:
            raise RuntimeError("hipMalloc failed")

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        if DeviceArray._hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy failed")
        return out

    def free(self):
        if self.p:
            DeviceArray._hip.hipFree(self.p)
            self.p = None


def run_path(g, mats, precision, steps, warmup, ctx, profile=False, **kw):
    """Steps of one solve from a cold start plus the bond checks on the resident solution into device
    arrays (phi, mode in HBM); the same checks into host arrays, with the copies over PCIe, are timed
    apart (checks_host_ms). profile: after the timed steps, as many solves more with the library timing
    its dominant kernels live (the roofline's figures; a timed launch costs a few microseconds more,
    so these solves are not the timed ones)."""
    from opus_destruct_2_amd import stress
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx)
    t = time.perf_counter()
    assert p.assemble(stress.default_options(precision=precision, **{**AMG_KW, **kw}))
    setup_ms = (time.perf_counter() - t) * 1e3
    info = p.info()
    log(f"  precision={precision} setup {setup_ms:.1f} ms levels {info['level_rows']} blocks {info['level_blocks']}")
    dphi, dmode = DeviceArray(8 * g.m), DeviceArray(g.m)

    def checks_on_device():
        st = stress.lib().svxg_problem_bond_checks_noalloc(p.h, None, 1.0, dphi.p, dmode.p)
        if st != 0:
            raise RuntimeError(f"bond checks on the device: status {st}: {stress.lib().svxg_last_error()}")

    for _ in range(warmup):
        u, r = p.solve(g.f, rtol=1e-8, maxit=5000)
        checks_on_device()
    dev_ms, wall_ms, chk_ms, iters = [], [], [], []
    for s in range(steps):
        t0 = time.perf_counter()
        u, r = p.solve(g.f, rtol=1e-8, maxit=5000)
        t1 = time.perf_counter()
        checks_on_device()
        t2 = time.perf_counter()
        assert r["converged"], r
        dev_ms.append(p.last_timing()["solve_ms"])
        wall_ms.append((t1 - t0) * 1e3)
        chk_ms.append((t2 - t1) * 1e3)
        iters.append(r["iters"])
        log(f"  step {s}: solve {dev_ms[-1]:.2f} ms device ({wall_ms[-1]:.2f} wall), checks {chk_ms[-1]:.2f} ms, "
            f"iters {r['iters']} refinements {r['refinements']} rel_res {r['rel_res']:.2e}")
    # (the live kernel timing runs on further solves of its own, so that it does not slow the timed ones)
    kstats = None
    if profile:
        ctx.set_profiling(True)
        for k in range(4):
            ctx.kernel_stats(k)
        for s in range(steps):
            p.solve(g.f, rtol=1e-8, maxit=5000)
        ctx.set_profiling(False)
        kstats = {k: ctx.kernel_stats(k) for k in range(4)}
    # the same checks into host arrays (phi and modes over PCIe), and the device arrays' bits against them
    out = (np.zeros(g.m), np.zeros(g.m, np.uint8))
    host_ms = []
    for _ in range(max(1, steps)):
        t1 = time.perf_counter()
        phi, mode = p.bond_checks(None, out=out)
        host_ms.append((time.perf_counter() - t1) * 1e3)
    if not (np.array_equal(dphi.get(np.float64, g.m).view(np.uint64), phi.view(np.uint64))
            and np.array_equal(dmode.get(np.uint8, g.m), mode)):
        raise RuntimeError("bond checks: the device arrays differ from the host arrays")
    dphi.free()
    dmode.free()
    out = {
        "setup_ms": setup_ms, "solve_ms": float(np.mean(dev_ms)), "wall_ms": float(np.mean(wall_ms)),
        "checks_ms": float(np.mean(chk_ms)), "checks_host_ms": float(np.median(host_ms)),
        "iters": int(np.median(iters)), "refinements": r["refinements"],
        "rel_res": r["rel_res"], "u": u, "phi": phi, "kstats": kstats, "info": info,
    }
    p.close()
    return out


def roofline(res, precision, shape):
    names = {0: "k_pcg_spmv<double>", 1: "k_pcg_spmv<float>", 2: "k_cheb_step<double, 1, false> (level 0)",
             3: "k_cheb_step<float, 1, false> (level 0" + (", FP16 block copies)" if res["info"]["vcycle_fp16"] else ")")}
    best = max(range(4), key=lambda k: res["kstats"][k][1])  # (the kind with the most device time)
    n, ms, byts = res["kstats"][best]
    if n == 0 or ms <= 0:
        return None
    avg_s = ms / n / 1e3
    per_launch = byts / n
    achieved = per_launch / avg_s / 1e9
    traffic = MEASURED_TRAFFIC.get((names[best],) + tuple(shape))
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": names[best],
            "launches": n, "avg_us": round(avg_s * 1e6, 2), "bytes_per_launch": int(per_launch)}


def solve_bytes(info, mixed, iters, refinements, cheb_degree=2, coarse_degree=1):
    """Algorithmic HBM bytes of one solve (each array a kernel reads or writes counted once a launch,
    gathered x once): per PCG iteration the FP64 K product, the PCG's vector updates and the V-cycle
    of every level (its Chebyshev steps, residual, restriction and prolongation over the level's
    block storage: FP16 blocks of 80 bytes, else 36 values of the cycle's precision, 4 bytes of column
    index a block; vectors of the cycle's precision), the dense coarsest solve (its FP64 inverse);
    plus an FP64 residual a refinement. The formula of the path's algorithmic bytes per unit of work."""
    rows, blocks, pblocks = info["level_rows"], info["level_blocks"], info["level_pblocks"]
    L = len(rows)
    w = 4 if mixed else 8                      # the V-cycle's vector and arithmetic precision
    blk = 80 if info["vcycle_fp16"] else 36 * w  # a stored block
    n0, nb0 = rows[0], blocks[0]
    k_pass = nb0 * (288 + 4) + 4 * (n0 + 1)
    per_it = k_pass + 2 * 6 * n0 * 8                       # q = K p: p gathered once, q written
    per_it += 6 * n0 * (6 * 8 + (4 if mixed else 0))       # x += a p, r -= a q (x, p, r, q read; x, r written; FP32 r)
    per_it += 6 * n0 * ((4 if mixed else 8) + 8 + 8)       # p = z + b p
    per_it += 6 * n0 * 8                                   # r read again for r . z
    dense = info.get("dense_rows", 0)
    for l in range(L):
        n, nb = rows[l], blocks[l]
        terminal = l == L - 1
        if terminal:
            if dense:
                per_it += 8 * dense * dense + 2 * dense * w
            continue
        dg = cheb_degree if l == 0 else coarse_degree
        nc, npb = rows[l + 1], pblocks[l]
        mat = nb * (blk + 4) + 4 * (n + 1)
        step = mat + blk * n + 5 * 6 * n * w                # r = b - A x, d = c1 d + c2 D^-1 r, x += d
        fresh = blk * n + 4 * 6 * n * w                     # from x = 0: b, D^-1 b, d and x written
        per_it += fresh + (2 * dg - 1) * step
        per_it += mat + 3 * 6 * n * w                       # residual
        per_it += npb * (blk + 4) + 4 * (nc + 1) + 6 * n * w + 6 * nc * w  # restriction (transposed copies)
        per_it += npb * (blk + 4) + 4 * (n + 1) + 6 * nc * w + 2 * 6 * n * w  # prolongation
    return iters * per_it + refinements * (k_pass + 3 * 6 * n0 * 8)


def copy_bandwidth(mib=1024):
    """A plain device copy's rate (bytes read + written / time, best of five) measured in a child
    process before this one touches the GPU (torch's HIP runtime and the library's cannot share a
    process): the HBM rate a streaming kernel can reach here, beside the 8 TB/s peak."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--copy-probe", str(mib)], capture_output=True,
                       text=True, timeout=300)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1]).get("copy_gbs")


def cpu_share():
    """(threads, facts): the host threads a CPU baseline uses (the CPUs this process may run on, capped
    by the cgroup's CPU quota and OMP_NUM_THREADS where either is set, as on a machine whose GPU owns a
    share of the host), with the CPU model, nproc and that share."""
    nproc = os.cpu_count() or 1
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            share = min(share, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return share, {"cpu_model": model, "nproc": nproc, "threads_all": share}


def cpu_solve(graphs, mats, threads, runs, opts=None, keep=False):
    """The reference's own solve (oracle/_ref, compiled from its sources; the oracle's restatement,
    one thread, where _ref is absent) of `graphs` one after another, to 1e-8: `runs` runs of assemble
    + cold solve; medians of the runs' setup and solve seconds (summed over the graphs)."""
    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    if kind == "reference":
        oracle.set_ref_threads(threads)
    else:
        threads = 1
    setups, solves, iters, out = [], [], [], []
    for run in range(runs):
        su = so = 0.0
        its = []
        for g in graphs:
            p = oracle.CpuProblem(g.nodes, g.bonds, mats, "ref" if kind == "reference" else "oracle")
            t = time.perf_counter()
            assert p.assemble(opts or oracle.reference_options())
            t1 = time.perf_counter()
            u, r = p.solve(g.f, rtol=1e-8, maxit=5000)
            t2 = time.perf_counter()
            assert r["converged"], r
            su += t1 - t
            so += t2 - t1
            its.append(r["iters"])
            if keep and run == runs - 1:
                phi, _ = p.bond_checks(u)
                out.append((u, phi))
        setups.append(su)
        solves.append(so)
        iters = its
    dofs = sum(g.dofs for g in graphs)
    solve_s = float(np.median(solves))
    return {"kind": kind, "threads": threads, "runs": runs, "setup_s": round(float(np.median(setups)), 4),
            "solve_s": round(solve_s, 4), "dof_per_s": round(dofs / solve_s, 1),
            "iters": int(max(iters)), "solutions": out}


def compare(u, phi, u_ref, phi_ref, band=1e-4):
    """Errors of one solution against a reference one: u relative to max |u_ref|, phi relative to
    max(1, |phi_ref|), and the failed bonds {phi >= 1}, which must agree except within `band` of 1."""
    a, b = phi >= 1.0, phi_ref >= 1.0
    near = (np.abs(phi - 1.0) <= band) | (np.abs(phi_ref - 1.0) <= band)
    bad = int(((a != b) & ~near).sum())
    return {"u_rel": float(np.abs(u - u_ref).max() / max(np.abs(u_ref).max(), 1e-300)),
            "phi_rel": float((np.abs(phi - phi_ref) / np.maximum(1.0, np.abs(phi_ref))).max()),
            "failed": int(a.sum()), "failed_ref": int(b.sum()), "failed_sets_agree": bad == 0,
            "failed_disagreeing_outside_1e-4": bad}


def merge_compare(recs):
    """compare() records of a sample merged: the largest errors, the failed counts summed."""
    if not recs:
        return None
    return {"u_rel": max(r["u_rel"] for r in recs), "phi_rel": max(r["phi_rel"] for r in recs),
            "failed": sum(r["failed"] for r in recs), "failed_ref": sum(r["failed_ref"] for r in recs),
            "failed_sets_agree": all(r["failed_sets_agree"] for r in recs),
            "failed_disagreeing_outside_1e-4": sum(r["failed_disagreeing_outside_1e-4"] for r in recs),
            "graphs": len(recs)}


def cpu_baseline(name, graphs, mats, runs_all, runs_one, opts=None, keep=False):
    """BASELINE's CPU protocol on one workload: the reference's solve with every host thread of this
    process's share and with one thread, medians of the runs, setup and solve apart."""
    threads, facts = cpu_share()
    t = time.perf_counter()
    allc = cpu_solve(graphs, mats, threads, runs_all, opts, keep=keep)
    one = cpu_solve(graphs, mats, 1, runs_one, opts) if runs_one > 0 else None
    sols = allc.pop("solutions")
    if one is not None:
        one.pop("solutions")
    rec = {"workload": name, "graphs": len(graphs), "dofs": int(sum(g.dofs for g in graphs)), **facts,
           "all_threads": allc, "one_thread": one, "wall_s": round(time.perf_counter() - t, 1)}
    log(f"  cpu baseline {name}: {rec}")
    return rec, sols


def run_batch(mats, ctx, count, steps, sample_plain, sigma=0.75, cpu_runs=(0, 0)):
    """Batched uneven graphs: `count` pieces (log-normal sizes, median 200, spread sigma, 16..20000
    nodes; even ones on supports, odd ones pinned), one batch, mixed precision to 1e-8; against the
    plain FP64 path solving graphs one at a time (a sample). On the sample: the batch's answers
    against the plain path's and, with cpu_runs, against the reference's CPU solve of the same pieces
    (the piece path: multigrid at every size), which is also timed."""
    import oracle
    from opus_destruct_2_amd import gen, stress
    from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED
    sizes = gen.piece_sizes(count, sigma=sigma, seed=3)
    gs = [gen.piece(int(n), pinned=bool(k & 1), seed=5000 + k) for k, n in enumerate(sizes)]
    dofs = sum(g.dofs for g in gs)
    log(f"batched: {count} graphs, {sum(g.n for g in gs)} nodes, {dofs} DOFs (sizes {sizes.min()}..{sizes.max()}, "
        f"median {int(np.median(sizes))}, sigma {sigma})")
    opts = stress.default_options(precision=SVXG_MIXED, amg_min_nodes=0, **AMG_KW)
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    t = time.perf_counter()
    ok = stress.assemble_batch(probs)
    setup_ms = (time.perf_counter() - t) * 1e3
    assert all(ok)
    fs = [g.f for g in gs]
    us, rs = stress.solve_batch(probs, fs, rtol=1e-8)
    dev, wall = [], []
    for s_ in range(steps):
        t = time.perf_counter()
        us, rs = stress.solve_batch(probs, fs, rtol=1e-8)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(probs[0].last_timing()["solve_ms"])
        assert all(r["converged"] for r in rs)
        log(f"  batch step {s_}: {dev[-1]:.2f} ms device ({wall[-1]:.2f} wall), iters max "
            f"{max(r['iters'] for r in rs)} median {int(np.median([r['iters'] for r in rs]))}")
    solve_ms = float(np.median(dev))
    fp16 = sum(1 for p in probs if p.info()["vcycle_fp16"])
    lean = sum(1 for p in probs if p.info()["lean"])
    # the plain path: FP64, one graph at a time; the batch's answers on the same graphs against its
    pick = list(range(0, count, max(1, count // sample_plain)))[:sample_plain]
    pdofs, pms = 0, 0.0
    popt = stress.default_options(precision=SVXG_FP64, amg_min_nodes=0, **AMG_KW)
    plain_sol, vs_plain = [], []
    for k in pick:
        g = gs[k]
        p = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=popt)
        assert p.assemble()
        p.solve(g.f, rtol=1e-8)
        u, r = p.solve(g.f, rtol=1e-8)
        assert r["converged"]
        pdofs += g.dofs
        pms += p.last_timing()["solve_ms"]
        phi, _ = p.bond_checks(u)
        phib, _ = probs[k].bond_checks(us[k])
        plain_sol.append((u, phi, phib))
        vs_plain.append(compare(us[k], phib, u, phi))
        p.close()
    plain = pdofs / (pms / 1e3)
    batched = dofs / (solve_ms / 1e3)
    for p in probs:
        p.close()
    rec = {"graphs": count, "dofs": dofs, "sizes": [int(sizes.min()), int(np.median(sizes)), int(sizes.max())],
           "sigma": sigma, "graphs_fp16_vcycle": fp16, "graphs_lean": lean,
           "value": round(batched, 1), "unit": "DOF/s", "solve_ms": round(solve_ms, 3),
           "wall_ms_pcie_inclusive": round(float(np.median(wall)), 3), "setup_ms": round(setup_ms, 1),
           "precision": "mixed", "iters_max": max(r["iters"] for r in rs),
           "plain_fp64_single_graph": {"value": round(plain, 1), "sample_graphs": len(pick), "sample_dofs": pdofs,
                                       "solve_ms_sum": round(pms, 3)},
           "batched_vs_plain": round(batched / plain, 2),
           "batched_mixed_vs_plain_fp64": merge_compare(vs_plain)}
    cpu = None
    if cpu_runs[0] > 0:
        sample = [gs[k] for k in pick]
        cpu, sols = cpu_baseline("batched pieces: the plain sample", sample, mats, cpu_runs[0], cpu_runs[1],
                                 oracle.reference_options(amg_min_nodes=0), keep=True)
        rec["error_vs_cpu_reference"] = {
            "plain_fp64": merge_compare([compare(pu, pp, cu, cp) for (pu, pp, _), (cu, cp) in zip(plain_sol, sols)]),
            "batched_mixed": merge_compare([compare(us[k], pb, cu, cp)
                                            for k, (_, _, pb), (cu, cp) in zip(pick, plain_sol, sols)])}
    return rec, cpu


def run_cascade(mats, ctx, g, target, cpu=None):
    """The break-round cascade of the single large structure (the block of the headline): its
    strengths scaled so that the largest phi under gravity is `target`, then solve, judge, break and
    detach round after round until nothing is over strength (svxg_break_rounds_batch, mixed
    precision to 1e-8), K, the multigrid and the solution staying on the device. The judge alone is
    timed on the resident solution; cpu: the reference's solve of the same block (setup and solve
    seconds) for a per-round CPU figure (its setup, as the device's rebuilds, is counted once a round)."""
    import oracle
    from opus_destruct_2_amd import stress
    from opus_destruct_2_amd.abi import SVXG_MIXED
    opts = stress.default_options(precision=SVXG_MIXED, **AMG_KW)
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts)
    assert p.assemble()
    u, r = p.solve(g.f, rtol=1e-8)
    phi, _ = p.bond_checks()
    jo = stress.default_judge_options(fragility=target / float(phi.max()))
    judge_ms = []
    for _ in range(5):
        t = time.perf_counter()
        j = stress.judge_batch([p], [{}], jo)[0]
        judge_ms.append((time.perf_counter() - t) * 1e3)
    p.close()
    p = stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts)
    assert p.assemble()
    t = time.perf_counter()
    us, res = stress.break_rounds_batch([p], [g.f], None, jo, rtol=1e-8)
    wall = (time.perf_counter() - t) * 1e3
    c = res[0]
    p.close()
    rec = {"target_phi": target, "end": c["end"], "rounds": c["rounds"], "breaks": int(len(c["breaks"])),
           "retired": int(len(c["retired"])), "solves": c["solves"], "iters": c["iters"], "rebuilds": c["rebuilds"],
           "wall_ms": round(wall, 1), "ms_per_round": round(wall / max(1, c["rounds"] + 1), 2),
           "first_judgement": {"over": j["over"], "picked": j["picked"],
                               "wall_ms_median": round(float(np.median(judge_ms)), 3)}}
    if cpu is not None:
        # the oracle's judge of the reference's solution, timed: with the reference's setup and
        # solve it is what one round costs on the host cores
        q = oracle.CpuProblem(g.nodes, g.bonds, mats)
        t = time.perf_counter()
        oracle.CpuJudge(q).judge(cpu["u"], jo, {})
        js = time.perf_counter() - t
        per = cpu["setup_s"] + cpu["solve_s"] + js
        rec["cpu_round_s"] = {"setup": round(cpu["setup_s"], 3), "solve": round(cpu["solve_s"], 3),
                              "judge": round(js, 3), "total": round(per, 3), "cores": cpu["cores"],
                              "kind": cpu["kind"]}
        rec["vs_cpu_per_round"] = round(per * 1e3 / rec["ms_per_round"], 1)
    log(f"  cascade: {rec}")
    return rec


def run_graph(mats, ctx, shape, cpu=True):
    """The graph of a voxel structure: a Voronoi block of `shape` voxels (fragments of about 2^3
    voxels, 15 % of them void, on an anchored ground layer) extracted on the host into nodes and face
    lists (opus_destruct_2_amd/voxels.py), the bonds' sections computed on the device
    (svxg_bond_sections, wall time with the copies), against the oracle's restatement on one host
    thread; the whole graph on the device (svxg_lattice_graph: two calls, the counts then the
    records, wall time with the copies); then the graph assembled and solved (mixed, 1e-8)."""
    import oracle
    from opus_destruct_2_amd import stress, voxels
    from opus_destruct_2_amd.abi import SVXG_MIXED
    h = 0.25
    t = time.perf_counter()
    vox, frag, lo = voxels.voronoi_block(*shape, cell=2, voids=0.15, seed=1)
    gen_s = time.perf_counter() - t
    t = time.perf_counter()
    nodes, heads, faces = voxels.extract(vox, frag, lo, h, mats)
    extract_s = time.perf_counter() - t
    walls = []
    for _ in range(3):
        t = time.perf_counter()
        bonds = voxels.bond_sections(nodes, heads, faces, vox, lo, h, mats, ctx=ctx)
        walls.append((time.perf_counter() - t) * 1e3)
    rec = {"voxels": int(vox.size), "nodes": int(len(nodes)), "bonds": int(len(heads)), "faces": int(len(faces)),
           "supports": int((heads["b"] < 0).sum()), "host_generate_s": round(gen_s, 2),
           "host_extract_s": round(extract_s, 2), "sections_wall_ms": round(float(np.median(walls)), 2),
           "bonds_per_s_wall": round(len(heads) / (float(np.median(walls)) / 1e3), 1)}
    # the whole graph on the device (faces, bonds, nodes, sections): svxg_lattice_graph
    lw = []
    for _ in range(3):
        t = time.perf_counter()
        dn, dh, df, db = voxels.lattice_graph(vox, frag, lo, h, mats, ctx=ctx)
        lw.append((time.perf_counter() - t) * 1e3)
    rec["device_graph_wall_ms"] = round(float(np.median(lw)), 2)
    rec["device_graph_equal_host"] = bool(dn.tobytes() == nodes.tobytes() and dh.tobytes() == heads.tobytes()
                                          and df.tobytes() == faces.tobytes() and db.tobytes() == bonds.tobytes())
    if cpu:
        t = time.perf_counter()
        ob = oracle.bond_sections(nodes, heads, faces, vox, lo, h, mats)
        cs = time.perf_counter() - t
        rec["cpu_sections_s"] = round(cs, 3)
        rec["cpu_cores"] = 1
        rec["bitwise_equal_cpu"] = bool(ob.tobytes() == bonds.tobytes())
    p = stress.StressProblem(nodes, bonds, mats, ctx, options=stress.default_options(precision=SVXG_MIXED, **AMG_KW))
    t = time.perf_counter()
    assert p.assemble()
    rec["setup_ms"] = round((time.perf_counter() - t) * 1e3, 1)
    u, r = p.solve(voxels.gravity(nodes), rtol=1e-8)
    u, r = p.solve(voxels.gravity(nodes), rtol=1e-8)
    rec["solve_ms"] = round(p.last_timing()["solve_ms"], 3)
    rec["iters"] = r["iters"]
    rec["dofs"] = int(6 * p.info()["nfree"])
    p.close()
    rec["extraction"] = run_extract(mats, ctx, vox, frag, lo, h, cpu=cpu)
    log(f"  graph from voxels: {rec}")
    return rec


def run_extract(mats, ctx, vox, frag, lo, h, cpu=True):
    """The structure the walk from a seed takes (svxg_extract_structure: World::Impl::extract and
    append_nodes on the device) over the same block, its fragments cut at chunk boundaries: with the
    reference's limits (60000 members, clusters of 16-voxel cells past 15000) and with none (every
    fragment a node). Wall time of the two calls (the counts, then the records) with the copies,
    against the oracle's sequential walk on one host thread, and the bits compared."""
    import oracle
    from opus_destruct_2_amd import voxels
    from opus_destruct_2_amd.abi import extract_options
    frag = voxels.chunk_fragments(frag, lo)
    fr = voxels.fragments(vox, frag, lo, h, mats)
    nz, ny, nx = frag.shape
    seed = int(frag[min(3, nz - 1), ny // 2, nx // 2])
    out = {"fragments": int(len(fr)), "seed": seed}
    for name, kw in (("limits", {}), ("whole", {"max_nodes": 1 << 20, "cluster_nodes": 1 << 20, "max_radius": 1e9})):
        opts = extract_options(**kw)
        walls = []
        for _ in range(3):
            t = time.perf_counter()
            d = voxels.extract_structure(vox, frag, lo, h, fr, seed, mats, opts=opts, ctx=ctx)
            walls.append((time.perf_counter() - t) * 1e3)
        r = {"members": int(len(d["members"])), "nodes": int(len(d["nodes"])), "bonds": int(len(d["bonds"])),
             "faces": int(len(d["faces"])), "cell": d["cell"], "truncated": d["truncated"],
             "device_wall_ms": round(float(np.median(walls)), 2),
             "members_per_s": round(len(d["members"]) / (float(np.median(walls)) / 1e3), 1)}
        if cpu:
            t = time.perf_counter()
            o = oracle.extract_structure(vox, frag, lo, h, fr, seed, mats, opts=opts)
            r["cpu_walk_s"] = round(time.perf_counter() - t, 3)
            r["cpu_cores"] = 1
            r["bitwise_equal_cpu"] = bool(all(o[k].tobytes() == d[k].tobytes() for k in
                                              ("members", "member_node", "nodes", "node_ident", "heads", "faces",
                                               "bonds", "bond_ident")))
        out[name] = r
    return out


class _Piece:
    """A bare problem handle for the serving driver (created straight through the C ABI)."""
    __slots__ = ("h", "n", "g", "broken")

    def __init__(self, g, mats, ctx):
        import ctypes
        from opus_destruct_2_amd import stress
        h = ctypes.c_void_p()
        st = stress.lib().svxg_problem_create(ctx.h, g.nodes.ctypes.data, g.n, g.bonds.ctypes.data, len(g.bonds),
                                              mats.ctypes.data, len(mats), ctypes.byref(h))
        assert st == 0
        self.h, self.n, self.g = h, g.n, g


def serving_sample(mats, ctx, pool, count, cpu_runs, seed=13):
    """A sample of served requests against the reference's CPU solve (the piece path): `count` pool
    graphs served cold, then each served again warm after losing 1-5 % of its bonds in place (as the
    stream's damaged residents do); the reference solves the same graphs, the second time with those
    bonds broken. Reports the served answers' u and phi errors and the failed-bond agreement."""
    import oracle
    from opus_destruct_2_amd import gen, serve, stress
    from opus_destruct_2_amd.abi import SVXG_MIXED
    rng = np.random.default_rng(seed)
    graphs = [pool[k] for k in range(0, len(pool), max(1, len(pool) // count))][:count]
    srv = serve.Server(ctx, precision=SVXG_MIXED, rtol=1e-8, maxit=5000, amg_min_nodes=0, window=4)
    pieces = [_Piece(g, mats, ctx) for g in graphs]
    n = len(pieces)

    def served(recs, tickets):
        out = []
        for t, pc in zip(tickets, pieces):
            r = recs[t]
            assert r["status"] == 0 and r["result"]["converged"], r
            phi = np.zeros(len(pc.g.bonds))
            mode = np.zeros(len(pc.g.bonds), np.uint8)
            st = stress.lib().svxg_problem_bond_checks(pc.h, None, 1.0, phi.ctypes.data, mode.ctypes.data)
            assert st == 0
            out.append((r["u"], phi))
        return out

    tk = srv.submit_many(pieces, [g.f for g in graphs], [False] * n, [None] * n, want_u=True)
    cold = served(srv.wait(tk), tk)
    cuts, damaged = [], []
    for g in graphs:
        c = max(1, int(len(g.bonds) * rng.uniform(0.01, 0.05)))
        ids = gen.safe_cuts(g, np.zeros(len(g.bonds), np.uint8), c, int(rng.integers(1 << 62)))
        cuts.append(ids)
        b = g.bonds.copy()
        b["broken"][ids] = 1
        damaged.append(gen.Graph(g.nodes, b, g.f, g.name + "-damaged"))
    tk = srv.submit_many(pieces, [None] * n, [True] * n, cuts, want_u=True)
    warm = served(srv.wait(tk), tk)
    for pc in pieces:
        srv.release(pc)
    srv.close()
    opts = oracle.reference_options(amg_min_nodes=0)
    cpu, ref_cold = cpu_baseline("serving sample: new pieces", graphs, mats, cpu_runs, 0, opts, keep=True)
    _, ref_warm = cpu_baseline("serving sample: the same pieces damaged", damaged, mats, 1, 0, opts, keep=True)
    rec = {"graphs": n, "dofs": int(sum(g.dofs for g in graphs)),
           "cold_vs_cpu_reference": merge_compare([compare(u, ph, cu, cp) for (u, ph), (cu, cp) in zip(cold, ref_cold)]),
           "damaged_warm_vs_cpu_reference": merge_compare(
               [compare(u, ph, cu, cp) for (u, ph), (cu, cp) in zip(warm, ref_warm)])}
    return rec, cpu


def run_serving(mats, ctx, lams, ticks, warm_ticks, life, pool_size, seed=11, setup_workers=None, sample=0,
                cpu_runs=0, ticks_low=0):
    """Serving stream: 60 Hz ticks; each tick, Poisson(lam) new structures (sizes as the batched
    leg's, half on supports, half pinned; multigrid at every size, as pieces are) arrive and are solved cold (setup + solve), and 10 % of the
    resident structures that are idle lose 1-5 % of their bonds and are solved again, warm, on their
    resident load; a structure leaves `life` ticks after it arrived (its device memory returns to
    the budget). Cuts never leave a mechanism (each free node keeps a path to a support or its pin:
    a detached part would become a piece of its own). Mixed precision to 1e-8 through svxg_server (continuous batching, admission by the
    context's HBM budget). A tick's latency: from its first submission to its last completion.
    ticks_low: measured ticks at the lowest rate (so that its p99 is a percentile of hundreds of
    ticks, not the worst of 90); the others measure `ticks`."""
    from opus_destruct_2_amd import gen, serve
    from opus_destruct_2_amd.abi import SVXG_MIXED
    rng = np.random.default_rng(seed)
    sizes = gen.piece_sizes(pool_size, seed=seed)
    pool = [gen.piece(int(n), pinned=bool(k & 1), seed=9000 + k) for k, n in enumerate(sizes)]
    out = {"tick_ms": 1000.0 / 60.0, "life_ticks": life, "warmup_ticks": warm_ticks, "pool_graphs": pool_size,
           "per_lambda": []}
    period = 1.0 / 60.0
    for lam in lams:
        ticks_here = ticks_low if ticks_low > 0 and lam == min(lams) else ticks
        skw = {} if setup_workers is None else {"setup_workers": setup_workers}
        srv = serve.Server(ctx, precision=SVXG_MIXED, rtol=1e-8, maxit=5000, amg_min_nodes=0, window=4, **skw)
        residents = []  # [piece, birth tick, cut offset]
        busy = {}  # ticket -> (tick, piece index or None, dofs)
        in_flight = set()
        tick_first, tick_last, tick_dofs, tick_req = {}, {}, {}, {}
        iters, lat_req, unconverged = [], [], 0
        stage = {"setup": [], "queue": [], "solve": []}
        done_dofs, t_meas0, t_meas1 = 0, None, None
        t0 = time.perf_counter()

        def collect(wait_ms=0):
            nonlocal done_dofs, unconverged
            for t, rec in srv.poll(wait_ms=wait_ms).items():
                k, piece, dofs = busy.pop(t)
                in_flight.discard(id(piece))
                assert rec["status"] == 0, rec
                unconverged += 0 if rec["result"]["converged"] else 1
                tick_last[k] = max(tick_last.get(k, 0.0), rec["t_done_ms"])
                if k >= warm_ticks:
                    done_dofs += dofs
                    iters.append(rec["result"]["iters"])
                    lat_req.append(rec["t_done_ms"] - rec["t_submit_ms"])
                    stage["setup"].append(rec["t_ready_ms"] - rec["t_submit_ms"])
                    stage["queue"].append(rec["t_admit_ms"] - rec["t_ready_ms"])
                    stage["solve"].append(rec["t_done_ms"] - rec["t_admit_ms"])

        for k in range(warm_ticks + ticks_here):
            target = t0 + k * period
            while True:
                collect()
                left = target - time.perf_counter()
                if left <= 0:
                    break
                time.sleep(min(left, 0.001))
            if k == warm_ticks:
                t_meas0 = srv.now_ms()
            # departures
            keep = []
            for r in residents:
                if k - r[1] >= life and id(r[0]) not in in_flight:
                    srv.release(r[0])
                else:
                    keep.append(r)
            residents = keep
            # damage: 10 % of the idle residents lose 1-5 % of their bonds
            idle = [r for r in residents if id(r[0]) not in in_flight]
            hit = [idle[i] for i in rng.choice(len(idle), size=len(idle) // 10, replace=False)] if idle else []
            probs, fs, warm, rem = [], [], [], []
            for r in hit:
                piece = r[0]
                m = len(piece.g.bonds)
                c = max(1, int(m * rng.uniform(0.01, 0.05)))
                ids = gen.safe_cuts(piece.g, piece.broken, c, int(rng.integers(1 << 62)))
                piece.broken[ids] = 1
                probs.append(piece)
                fs.append(None)
                warm.append(True)
                rem.append(ids)
            # arrivals
            for _ in range(rng.poisson(lam)):
                gi = int(rng.integers(pool_size))
                piece = _Piece(pool[gi], mats, ctx)
                piece.broken = np.zeros(len(pool[gi].bonds), np.uint8)
                residents.append([piece, k, 0])
                probs.append(piece)
                fs.append(pool[gi].f)
                warm.append(False)
                rem.append(None)
            if not probs:
                continue
            t_sub = srv.now_ms()
            tk = srv.submit_many(probs, fs, warm, rem, want_u=False)
            tick_first[k] = t_sub
            tick_req[k] = len(probs)
            tick_dofs[k] = sum(6 * p.n for p in probs)
            for t, piece in zip(tk, probs):
                busy[t] = (k, piece, 6 * piece.n)
                in_flight.add(id(piece))
        t_drain = time.perf_counter()
        while busy and time.perf_counter() - t_drain < 120.0:
            collect(wait_ms=20)
        if busy:
            raise RuntimeError(f"serving lambda={lam}: {len(busy)} requests still in flight after 120 s")
        st = srv.stats()
        for r in residents:
            srv.release(r[0])
        srv.close()
        lat = np.array([tick_last[k] - tick_first[k] for k in range(warm_ticks, warm_ticks + ticks_here) if k in tick_first])
        wall_s = (max(tick_last.values()) - t_meas0) / 1e3
        rec = {"lambda": lam, "measured_ticks": ticks_here, "p50_tick_ms": round(float(np.percentile(lat, 50)), 2),
               "p99_tick_ms": round(float(np.percentile(lat, 99)), 2), "max_tick_ms": round(float(lat.max()), 2),
               "ticks_over_16.7ms": int((lat > 1000.0 / 60.0).sum()),
               "requests_per_tick": round(float(np.mean([tick_req[k] for k in tick_req if k >= warm_ticks])), 1),
               "dofs_per_tick": int(np.mean([tick_dofs[k] for k in tick_dofs if k >= warm_ticks])),
               "value": round(done_dofs / wall_s, 1), "unit": "DOF/s",
               "p50_request_ms": round(float(np.percentile(lat_req, 50)), 2),
               "p99_request_ms": round(float(np.percentile(lat_req, 99)), 2),
               "stage_p50_ms": {k_: round(float(np.median(v)), 2) for k_, v in stage.items()},
               "stage_p99_ms": {k_: round(float(np.percentile(v, 99)), 2) for k_, v in stage.items()},
               "iters_median": int(np.median(iters)), "iters_max": int(max(iters)), "unconverged": unconverged,
               "server": {k_: (round(v, 1) if isinstance(v, float) else v) for k_, v in st.items()}}
        log(f"  serving lambda={lam}: p50 {rec['p50_tick_ms']} ms, p99 {rec['p99_tick_ms']} ms per tick, "
            f"{rec['requests_per_tick']} requests / tick, {rec['value'] / 1e6:.1f}M DOF/s, "
            f"{rec['ticks_over_16.7ms']} of {len(lat)} ticks over 16.7 ms; iters median {rec['iters_median']} "
            f"max {rec['iters_max']}; stages p50 {rec['stage_p50_ms']} p99 {rec['stage_p99_ms']}; server {rec['server']}")
        out["per_lambda"].append(rec)
    if sample > 0:
        out["sample"], out["sample_cpu"] = serving_sample(mats, ctx, pool, sample, max(1, cpu_runs))
        log(f"  serving sample: {out['sample']}")
    return out


def run_city(mats, ctx, count, rounds, rate, seed=21):
    """City chunk: `count` resident buildings (towers on supports, 5k..150k nodes, log-normal sizes
    with median 20k, one material a storey) solved together, one batch a round, mixed precision to
    1e-8 from a cold start under gravity and a wind load that turns every round; the multigrids stay
    resident. Collapse events arrive at `rate` per second of the timed rounds' wall time: each cuts
    one building's ground-storey columns under half its footprint (bonds out in place) and re-sets
    up that building's multigrid at the start of the next round, inside its timed region. After the
    last round a 5 % sample (every 20th building) is solved on the FP64 path and compared."""
    from opus_destruct_2_amd import gen, stress
    from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED
    t = time.perf_counter()
    shapes = gen.city_shapes(count, seed=seed)
    gs = [gen.building(*s, seed=seed * 1000 + k) for k, s in enumerate(shapes)]
    gen_s = time.perf_counter() - t
    nodes, dofs = sum(g.n for g in gs), sum(g.dofs for g in gs)
    sizes = np.array([g.n for g in gs])
    log(f"city: {count} buildings, {nodes} nodes, {dofs} DOFs (sizes {sizes.min()}..{sizes.max()}, median "
        f"{int(np.median(sizes))}), generated in {gen_s:.1f} s")
    opts = stress.default_options(precision=SVXG_MIXED, **AMG_KW)
    probs = [stress.StressProblem(g.nodes, g.bonds, mats, ctx, options=opts) for g in gs]
    t = time.perf_counter()
    assert all(stress.assemble_batch(probs))
    setup_ms = (time.perf_counter() - t) * 1e3
    hbm = ctx.bytes_in_use()
    log(f"  setup of all buildings: {setup_ms:.0f} ms, {hbm / 2**30:.1f} GiB resident")
    rng = np.random.default_rng(seed)
    order = [int(k) for k in rng.permutation(count)]
    angle = lambda r: 2.0 * np.pi * r / 8.0  # noqa: E731
    t = time.perf_counter()
    us, rs = stress.solve_batch(probs, [gen.wind_load(g, angle(0)) for g in gs], rtol=1e-8)
    assert all(r["converged"] for r in rs)
    due = rate * (time.perf_counter() - t)  # collapses that arrived during the untimed first round
    log(f"  first round (untimed): {probs[0].last_timing()['solve_ms']:.1f} ms device, iters max "
        f"{max(r['iters'] for r in rs)}")
    hit_all, per_round, t_sum = [], [], 0.0
    for r in range(1, rounds + 1):
        fs = [gen.wind_load(g, angle(r)) for g in gs]
        t0 = time.perf_counter()
        nev = min(int(due), count - len(hit_all))
        due -= nev
        hit = order[len(hit_all):len(hit_all) + nev]
        hit_all += hit
        for k in hit:
            probs[k].remove_bonds(gen.ground_storey_cuts(gs[k]))
        if hit:
            assert all(stress.assemble_batch([probs[k] for k in hit]))
        t1 = time.perf_counter()
        us, rs = stress.solve_batch(probs, fs, rtol=1e-8)
        t2 = time.perf_counter()
        assert all(r_["converged"] for r_ in rs)
        due += rate * (t2 - t0)
        t_sum += t2 - t0
        its = [r_["iters"] for r_ in rs]
        rec = {"collapses": nev, "resetup_ms": round((t1 - t0) * 1e3, 1), "solve_ms_device":
               round(probs[0].last_timing()["solve_ms"], 1), "wall_ms": round((t2 - t0) * 1e3, 1),
               "iters_max": max(its), "iters_median": int(np.median(its)),
               "iters_max_collapsed": max([its[k] for k in hit], default=0)}
        per_round.append(rec)
        log(f"  round {r}: {rec}")
    # the 5 % sample on the FP64 path, same bonds and the last round's loads
    pick = list(range(0, count, 20))
    p64 = []
    for k in pick:
        p64.append(stress.StressProblem(gs[k].nodes, probs[k].bonds, mats, ctx,
                                        options=stress.default_options(precision=SVXG_FP64, **AMG_KW)))
    assert all(stress.assemble_batch(p64))
    u64, r64 = stress.solve_batch(p64, [fs[k] for k in pick], rtol=1e-8)
    assert all(r_["converged"] for r_ in r64)
    u_err, phi_err, failed_m, failed_d, disagree = 0.0, 0.0, 0, 0, 0
    for j, k in enumerate(pick):
        scale = np.abs(u64[j]).max()
        u_err = max(u_err, float(np.abs(us[k] - u64[j]).max() / scale))
        phi, _ = probs[k].bond_checks(us[k])
        phi64, _ = p64[j].bond_checks(u64[j])
        phi_err = max(phi_err, float(np.abs(phi - phi64).max() / max(1.0, np.abs(phi64).max())))
        a, b = phi >= 1.0, phi64 >= 1.0
        near = (np.abs(phi - 1.0) <= 1e-4) | (np.abs(phi64 - 1.0) <= 1e-4)
        failed_m += int(a.sum())
        failed_d += int(b.sum())
        disagree += int(((a != b) & ~near).sum())
        p64[j].close()
    for p in probs:
        p.close()
    solve_dev = float(np.median([x["solve_ms_device"] for x in per_round]))
    return {"buildings": count, "nodes": nodes, "dofs": dofs, "sizes": [int(sizes.min()), int(np.median(sizes)),
                                                                         int(sizes.max())],
            "precision": "mixed", "rtol": 1e-8, "rounds": rounds, "collapse_rate_per_s": rate,
            "collapses": len(hit_all), "setup_ms": round(setup_ms, 1), "resident_gib": round(hbm / 2**30, 2),
            "value": round(dofs * rounds / t_sum, 1), "unit": "DOF/s (wall, re-setups and PCIe copies included)",
            "device_solve_dof_per_s": round(dofs / (solve_dev / 1e3), 1), "per_round": per_round,
            "fp64_sample": {"buildings": len(pick), "u_rel_err_max": u_err, "phi_rel_err_max": phi_err,
                            "failed_bonds_mixed": failed_m, "failed_bonds_fp64": failed_d,
                            "failed_disagreeing_outside_1e-4": disagree}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baseline: runs with every host thread (median)")
    ap.add_argument("--cpu-runs-one-thread", type=int, default=1,
                    help="CPU baseline of the large block with one thread: runs (median; the smaller workloads "
                         "always take --cpu-runs)")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=48)
    ap.add_argument("--batch", type=int, default=4096, help="graphs in the batched leg (0: skip)")
    ap.add_argument("--plain-sample", type=int, default=128)
    ap.add_argument("--batch-sigma", type=float, default=0.75, help="log-normal spread of the batched leg's sizes")
    ap.add_argument("--batch-wide-sigma", type=float, default=1.32,
                    help="log-normal spread of the batched leg's full-range variant (0: skip it)")
    ap.add_argument("--serve-lambdas", default="50,200,800", help="arrivals per tick of the serving leg ('' skips)")
    ap.add_argument("--serve-ticks", type=int, default=90)
    ap.add_argument("--serve-ticks-low", type=int, default=360, help="measured ticks at the lowest arrival rate")
    ap.add_argument("--serve-warmup", type=int, default=30)
    ap.add_argument("--serve-life", type=int, default=20)
    ap.add_argument("--serve-pool", type=int, default=1024)
    ap.add_argument("--serve-setup-workers", type=int, default=None)
    ap.add_argument("--city", type=int, default=400, help="buildings in the city-chunk leg (0: skip)")
    ap.add_argument("--city-rounds", type=int, default=3)
    ap.add_argument("--city-rate", type=float, default=20.0, help="collapse events per second")
    ap.add_argument("--cascade", type=float, default=3.0,
                    help="largest phi of the cascade leg's block under gravity (0: skip the leg)")
    ap.add_argument("--graph", default="128x128x96",
                    help="voxels of the graph-from-voxels leg ('' skips it)")
    ap.add_argument("--smooth-levels", type=int, default=None,
                    help="levels with a smoothed prolongator, from the finest (0: all; default: the library's)")
    ap.add_argument("--cheb-degree", type=int, default=None)
    ap.add_argument("--coarse-cheb-degree", type=int, default=None)
    ap.add_argument("--amg", default="", metavar="KEY=VALUE,...",
                    help="further multigrid options of every leg (svxg_amg_options / svxg_stress_options field names)")
    ap.add_argument("--mixed-fp64-rows", type=int, default=None,
                    help="mixed precision: graphs below this many free nodes take FP64 K products (default: the library's)")
    ap.add_argument("--resident-rows", type=int, default=-1,
                    help="graphs up to this many free nodes step in one workgroup each (-1: the library's default)")
    ap.add_argument("--autotune", type=int, default=1, choices=[0, 1],
                    help="tile shapes tuned online per level shape (1, the library's default) or the untuned "
                         "defaults (0: fixed launch shapes, e.g. for counter passes)")
    ap.add_argument("--step-graphs", type=int, default=None, choices=[0, 1, 2],
                    help="--only legs: record tiled solve steps as HIP graphs (0 never, 1 large batches, 2 always; "
                         "default: the library's)")
    ap.add_argument("--only", choices=["head", "batch", "serve", "city", "cascade", "graph", "cpu"], default=None,
                    help="profiling: run only this leg and print its record (not the benchmark line)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one plain device copy of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the stress solve here is measured on one MI355X's share; "
              "--gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        import torch
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        best = None
        for _ in range(5):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            dst.copy_(src)
            b.record()
            torch.cuda.synchronize()
            best = a.elapsed_time(b) if best is None else min(best, a.elapsed_time(b))
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n,
                          "copy_gbs": round(2 * n / (best / 1e3) / 1e9, 1)}))
        return

    copy_gbs = copy_bandwidth() if args.only is None else None  # (before this process touches the GPU)
    from opus_destruct_2_amd import gen, stress
    from opus_destruct_2_amd.abi import SVXG_FP64, SVXG_MIXED
    mats = gen.materials()
    g = gen.block(args.nx, args.ny, args.nz, 0.5, 0.15, seed=1)
    log(f"workload {g.name}: {g.n} nodes, {g.m} bonds, {g.dofs} DOFs")
    if args.smooth_levels is not None:
        AMG_KW["smooth_levels"] = args.smooth_levels
    if args.cheb_degree is not None:
        AMG_KW["cheb_degree"] = args.cheb_degree
    if args.coarse_cheb_degree is not None:
        AMG_KW["coarse_cheb_degree"] = args.coarse_cheb_degree
    if args.mixed_fp64_rows is not None:
        AMG_KW["mixed_fp64_rows"] = args.mixed_fp64_rows
    for kv in filter(None, args.amg.split(",")):
        k, v = kv.split("=")
        AMG_KW[k.strip()] = float(v) if "." in v or "e" in v else int(v)
    ctx = stress.default_context()
    ctx.set_autotune(bool(args.autotune))
    if args.resident_rows >= 0:
        ctx.set_resident_rows(args.resident_rows)
    if args.only is not None:
        if args.step_graphs is not None:
            ctx.set_step_graphs(args.step_graphs)
        if args.only == "head":  # (the headline solves alone, with their live roofline)
            prec = SVXG_MIXED if args.precision == "mixed" else SVXG_FP64
            head = run_path(g, mats, prec, args.steps, args.warmup, ctx, profile=True)
            sb = solve_bytes(head["info"], prec == SVXG_MIXED, head["iters"], head["refinements"])
            rec = {"solve_ms": round(head["solve_ms"], 3), "iters": head["iters"],
                   "checks_ms": round(head["checks_ms"], 3), "checks_host_ms": round(head["checks_host_ms"], 3),
                   "roofline": roofline(head, args.precision, (args.nx, args.ny, args.nz)),
                   "solve_bytes": int(sb), "solve_frac": round(sb / (head["solve_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                   "autotune": ctx.tuned()}
        elif args.only == "batch":
            rec, _ = run_batch(mats, ctx, args.batch, max(1, args.steps // 2), args.plain_sample, args.batch_sigma)
        elif args.only == "cpu":  # (the CPU protocol on the block: every thread and one thread)
            rec, _ = cpu_baseline(g.name, [g], mats, args.cpu_runs, args.cpu_runs_one_thread)
        elif args.only == "serve":
            rec = run_serving(mats, ctx, [int(x) for x in args.serve_lambdas.split(",") if x], args.serve_ticks,
                              args.serve_warmup, args.serve_life, args.serve_pool,
                              setup_workers=args.serve_setup_workers, ticks_low=args.serve_ticks_low)
        elif args.only == "cascade":
            rec = run_cascade(mats, ctx, g, args.cascade)
        elif args.only == "graph":
            rec = run_graph(mats, ctx, tuple(int(x) for x in args.graph.split("x")), cpu=not args.no_cpu)
        else:
            rec = run_city(mats, ctx, args.city, args.city_rounds, args.city_rate)
        print(json.dumps({"only": args.only, "record": rec}), flush=True)
        return
    head_prec = SVXG_MIXED if args.precision == "mixed" else SVXG_FP64
    other_prec = SVXG_FP64 if head_prec == SVXG_MIXED else SVXG_MIXED
    head = run_path(g, mats, head_prec, args.steps, args.warmup, ctx, profile=True)
    other = run_path(g, mats, other_prec, max(1, args.steps // 2), 1, ctx)
    plain = other if head_prec == SVXG_MIXED else head
    mixed_res = head if head_prec == SVXG_MIXED else other
    step_s = (head["solve_ms"] + head["checks_ms"]) / 1e3
    value = g.dofs / step_s
    rl = roofline(head, args.precision, (args.nx, args.ny, args.nz))
    if rl is not None:
        rl["copy_gbs"] = copy_gbs
        # the whole solve against the same roofline: its algorithmic bytes / its device time / peak
        sb = solve_bytes(head["info"], head_prec == SVXG_MIXED, head["iters"], head["refinements"])
        rl["solve_bytes"] = int(sb)
        rl["solve_achieved"] = round(sb / (head["solve_ms"] / 1e3) / 1e9, 1)
        rl["solve_frac"] = round(sb / (head["solve_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if head_prec == SVXG_MIXED else "f64", "data": "synthetic",
        "config": {"workload": f"single large structure: block {args.nx}x{args.ny}x{args.nz}, 15% voids, "
                               f"{g.n} nodes, {g.dofs} DOFs, AMG-PCG to 1e-8 + bond checks",
                   "precision": args.precision,
                   "precision_detail": (PRECISION_MIXED if head_prec == SVXG_MIXED else PRECISION_FP64)
                   + ("; V-cycle matrices, prolongators and inverse diagonal blocks stored as FP16 blocks "
                      "(computed in FP32)" if head["info"]["vcycle_fp16"] else ""),
                   "iters": head["iters"], "refinements": head["refinements"],
                   "setup_ms": round(head["setup_ms"], 1), "solve_ms": round(head["solve_ms"], 3),
                   "checks_ms": round(head["checks_ms"], 3), "checks_host_ms": round(head["checks_host_ms"], 3),
                   "wall_ms_pcie_inclusive": round(head["wall_ms"] + head["checks_host_ms"], 3)},
        "plain_fp64": {"value": round(g.dofs / ((plain["solve_ms"] + plain["checks_ms"]) / 1e3), 1),
                       "precision_detail": PRECISION_FP64,
                       "solve_ms": round(plain["solve_ms"], 3), "iters": plain["iters"],
                       "setup_ms": round(plain["setup_ms"], 1)},
        "roofline": rl,
    }
    line["mixed_vs_fp64"] = compare(mixed_res["u"], mixed_res["phi"], plain["u"], plain["phi"])
    # the other mixed split: FP32 K products too, inside FP64 iterative refinement
    ir = run_path(g, mats, SVXG_MIXED, max(1, args.steps // 2), 1, ctx, mixed_fp64_rows=0)
    line["mixed_fp32_products_ir"] = {
        "value": round(g.dofs / ((ir["solve_ms"] + ir["checks_ms"]) / 1e3), 1), "solve_ms": round(ir["solve_ms"], 3),
        "iters": ir["iters"], "refinements": ir["refinements"],
        "vs_fp64": compare(ir["u"], ir["phi"], plain["u"], plain["phi"])}
    # C1: the plain method on the small tower (8 x 8 x 20, FP64), and the mixed path beside it
    tw = gen.tower()
    c1 = run_path(tw, mats, SVXG_FP64, args.steps, args.warmup, ctx)
    c1m = run_path(tw, mats, SVXG_MIXED, args.steps, args.warmup, ctx)
    line["tower"] = {"workload": f"tower 8x8x20: {tw.n} nodes, {tw.dofs} DOFs", "dofs": tw.dofs,
                     "fp64": {"value": round(tw.dofs / ((c1["solve_ms"] + c1["checks_ms"]) / 1e3), 1),
                              "solve_ms": round(c1["solve_ms"], 3), "checks_ms": round(c1["checks_ms"], 3),
                              "iters": c1["iters"], "setup_ms": round(c1["setup_ms"], 1)},
                     "mixed": {"value": round(tw.dofs / ((c1m["solve_ms"] + c1m["checks_ms"]) / 1e3), 1),
                               "solve_ms": round(c1m["solve_ms"], 3), "iters": c1m["iters"],
                               "setup_ms": round(c1m["setup_ms"], 1)}}
    cpu_all = {}
    if args.batch > 0:
        line["batched"], cpu_all["batched_sample"] = run_batch(
            mats, ctx, args.batch, max(1, args.steps // 2), args.plain_sample, args.batch_sigma,
            (0, 0) if args.no_cpu else (args.cpu_runs, args.cpu_runs))
        # the same count of pieces over the configuration's whole 16..20,000-node range (a spread
        # whose largest of 4096 draws reaches the clamp): graphs of 16384+ free nodes among them take
        # the lean hierarchy inside the batch
        if args.batch_wide_sigma > 0:
            line["batched_full_range"], _ = run_batch(mats, ctx, args.batch, 1, min(8, args.plain_sample),
                                                      args.batch_wide_sigma)
    if args.serve_lambdas:
        lams = [int(x) for x in args.serve_lambdas.split(",") if x]
        log("serving: 60 Hz ticks, lambdas " + ", ".join(map(str, lams)))
        # (on the context the headline and the batch used: a long-lived server's state)
        line["serving"] = run_serving(mats, ctx, lams, args.serve_ticks, args.serve_warmup, args.serve_life,
                                      args.serve_pool, setup_workers=args.serve_setup_workers,
                                      sample=0 if args.no_cpu else 16, cpu_runs=args.cpu_runs,
                                      ticks_low=args.serve_ticks_low)
        if line["serving"].get("sample_cpu") is not None:
            cpu_all["serving_sample"] = line["serving"].pop("sample_cpu")
    if args.city > 0:
        cctx = stress.Context(0)  # (a context of its own, freed after the leg)
        line["city"] = run_city(mats, cctx, args.city, args.city_rounds, args.city_rate)
        cctx.close()
    cb = None
    if not args.no_cpu:
        log("cpu baseline: the reference's solver on the same workloads ...")
        c1cpu, s1 = cpu_baseline(tw.name, [tw], mats, args.cpu_runs, args.cpu_runs, keep=True)
        cpu_all["tower"] = c1cpu
        line["tower"]["error_vs_cpu_reference"] = {
            "fp64": compare(c1["u"], c1["phi"], *s1[0]), "mixed": compare(c1m["u"], c1m["phi"], *s1[0])}
        line["tower"]["fp64_vs_cpu_all_threads"] = round(
            (c1cpu["all_threads"]["solve_s"]) / (c1["solve_ms"] / 1e3), 1)
        c2cpu, s2 = cpu_baseline(g.name, [g], mats, args.cpu_runs, args.cpu_runs_one_thread, keep=True)
        cpu_all["block"] = c2cpu
        cu, cphi = s2[0]
        line["error_vs_cpu_reference"] = {"fp64": compare(plain["u"], plain["phi"], cu, cphi),
                                          "mixed": compare(mixed_res["u"], mixed_res["phi"], cu, cphi),
                                          "mixed_fp32_products_ir": compare(ir["u"], ir["phi"], cu, cphi)}
        allc = c2cpu["all_threads"]
        cb = {"value": allc["dof_per_s"], "unit": "DOF/s", "cores": allc["threads"], "kind": allc["kind"],
              "sample": f"the same block solved to 1e-8 ({allc['iters']} iterations), median of {allc['runs']} runs "
                        f"of assemble + cold solve on {allc['threads']} threads ({c2cpu['cpu_model']}, nproc "
                        f"{c2cpu['nproc']}): setup {allc['setup_s']:.2f} s, solve {allc['solve_s']:.2f} s"}
        if c2cpu["one_thread"]:
            cb["one_thread"] = {k: c2cpu["one_thread"][k] for k in ("dof_per_s", "setup_s", "solve_s", "runs")}
        cb["setup_s"], cb["solve_s"], cb["u"] = allc["setup_s"], allc["solve_s"], cu
    if args.cascade > 0:
        log(f"cascade: the block's break rounds, largest phi {args.cascade} under gravity")
        kctx = stress.Context(0)
        line["cascade"] = run_cascade(mats, kctx, g, args.cascade, cb)
        kctx.close()
    if args.graph:
        log(f"graph from voxels: a {args.graph} Voronoi block")
        gctx = stress.Context(0)
        line["graph_from_voxels"] = run_graph(mats, gctx, tuple(int(x) for x in args.graph.split("x")),
                                              cpu=not args.no_cpu)
        gctx.close()
    if cb is not None:
        for k in ("setup_s", "solve_s", "u"):
            cb.pop(k)
    line["cpu_baseline"] = cb
    line["cpu_baselines"] = cpu_all or None
    # the autotuner's choices on the headline's context (the block, the tower, the batch)
    line["autotune"] = ctx.tuned()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
