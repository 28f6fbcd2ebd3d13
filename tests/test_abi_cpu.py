"""The C ABI boundary without a GPU: libsvxg.so loads and exports every entry point that
include/svxg.h declares; the structs' layouts agree between the header and the Python mirrors;
the generator is deterministic (its graphs are fixed by their digests)."""
import ctypes
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from opus_destruct_2_amd import abi, gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "svxg.h")
LIB = os.path.join(ROOT, "opus_destruct_2_amd", "libsvxg.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|int64_t|double|const char\*)\s+(svxg_\w+)\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for must in ("svxg_ctx_create", "svxg_problem_create", "svxg_problem_assemble", "svxg_problem_solve",
                 "svxg_solve_batch", "svxg_problem_bond_checks", "svxg_problem_remove_bonds"):
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libsvxg.so not built"
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes_match_the_header():
    # sizes written in include/svxg.h's comments
    assert abi.NODE.itemsize == 40
    assert abi.BOND.itemsize == 216
    assert abi.MATERIAL.itemsize == 64
    assert ctypes.sizeof(abi.StressOptions) == 8 + 4 + 4 + ctypes.sizeof(abi.AmgOptions) + 8
    assert ctypes.sizeof(abi.PcgResult) == 24


def test_struct_layouts_match_a_c_compiler(tmp_path):
    """Every struct of include/svxg.h: size and field offsets as gcc lays them out = the ctypes mirror."""
    mirrors = {"svxg_amg_options": abi.AmgOptions, "svxg_stress_options": abi.StressOptions,
               "svxg_pcg_result": abi.PcgResult, "svxg_problem_info": abi.ProblemInfo,
               "svxg_server_options": abi.ServerOptions, "svxg_serve_done": abi.ServeDone,
               "svxg_server_stats": abi.ServerStats, "svxg_judge_options": abi.JudgeOptions,
               "svxg_judge_state": abi.JudgeState, "svxg_judge_result": abi.JudgeResult,
               "svxg_cascade_result": abi.CascadeResult, "svxg_voxel_box": abi.VoxelBox,
               "svxg_lattice": abi.Lattice, "svxg_extract_options": abi.ExtractOptions,
               "svxg_extract_result": abi.ExtractResult}
    dtypes = {"svxg_hinge_cut": abi.HINGE_CUT, "svxg_break": abi.BREAK, "svxg_bond_load": abi.BOND_LOAD,
              "svxg_section_head": abi.SECTION_HEAD, "svxg_section_face": abi.SECTION_FACE,
              "svxg_section_head_general": abi.SECTION_HEAD_GENERAL, "svxg_junction_sample": abi.JUNCTION_SAMPLE,
              "svxg_fragment": abi.FRAGMENT}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "svxg.h"', "int main(void) {"]
    for c, py in mirrors.items():
        lines.append(f'  printf("{c} %zu\\n", sizeof({c}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{c}.{f} %zu\\n", offsetof({c}, {f}));')
    for c, dt in dtypes.items():
        lines.append(f'  printf("{c} %zu\\n", sizeof({c}));')
        for f in dt.names:
            lines.append(f'  printf("{c}.{f} %zu\\n", offsetof({c}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for c, py in mirrors.items():
        assert int(got[c]) == ctypes.sizeof(py), c
        for f, _ in py._fields_:
            assert int(got[f"{c}.{f}"]) == getattr(py, f).offset, (c, f)
    for c, dt in dtypes.items():
        assert int(got[c]) == dt.itemsize, c
        for f in dt.names:
            assert int(got[f"{c}.{f}"]) == dt.fields[f][1], (c, f)


def test_default_options_without_a_device():
    from opus_destruct_2_amd import stress
    o = stress.default_options()
    assert o.rtol == 2e-3 and o.amg_min_nodes == 160 and o.amg.coarse_max == 48 and o.amg.strength == 0.1
    assert o.mixed_fp64_rows == 1 << 30 and o.vcycle_fp16_rows == 16384 and o.amg.coarse_cheb_degree == 1
    assert o.amg.lean_rows == 16384 and o.amg.lean_strength == 0.05
    j = stress.default_judge_options()
    assert (j.fragility, j.dif, j.break_band, j.quota_fraction, j.hinge_shape) == (1.0, 1.5, 0.85, 0.25, 1.3)
    assert (j.max_breaks, j.max_rounds, j.plastic_hinges, j.rule) == (256, 400, 1, 0)


def test_extract_default_options_without_a_device():
    """svxg_extract_default_options: WorldConfig's extraction defaults (world.hpp:142-148, 265), the
    same as abi.extract_options()."""
    from opus_destruct_2_amd import stress
    o = abi.ExtractOptions()
    stress.lib().svxg_extract_default_options(ctypes.byref(o))
    want = abi.extract_options()
    assert bytes(o) == bytes(want)
    assert (o.max_nodes, o.cluster_nodes, o.max_radius, o.cluster_cubes, o.outside) == (60000, 2500, 60.0, 0, 0)


def _digest(g):
    h = hashlib.sha256()
    h.update(g.nodes.tobytes())
    h.update(g.bonds.tobytes())
    h.update(g.f.tobytes())
    return h.hexdigest()[:16]


def test_generator_is_deterministic_and_fixed():
    a, b = gen.tower(), gen.tower()
    assert _digest(a) == _digest(b)
    assert a.n == 1280 and a.m == 3520 and a.dofs == 7680
    # the same bits on every machine: integer RNG and correctly rounded arithmetic only
    assert _digest(a) == FIXED["tower"]
    assert _digest(gen.piece(300, pinned=True, seed=5)) == FIXED["piece300p"]
    sizes = gen.piece_sizes(4096)
    assert sizes.min() >= 16 and sizes.max() <= 20000 and 150 <= np.median(sizes) <= 260


FIXED = {"tower": None, "piece300p": None}


@pytest.fixture(autouse=True, scope="module")
def _fixed_digests():
    path = os.path.join(ROOT, "tests", "golden", "generator_digests.txt")
    for line in open(path):
        k, v = line.split()
        FIXED[k] = v


def _grounded(g, broken):
    """Whether every node of g reaches a support bond or a pinned node over intact bonds."""
    parent = list(range(g.n + 1))  # node g.n: the ground

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for i in range(g.n):
        if g.nodes["fixed"][i]:
            parent[find(i)] = find(g.n)
    for k in range(len(g.bonds)):
        if broken[k]:
            continue
        a, b = int(g.bonds["a"][k]), int(g.bonds["b"][k])
        parent[find(a)] = find(g.n if b < 0 else b)
    return all(find(i) == find(g.n) for i in range(g.n))


def test_safe_cuts_never_leave_a_mechanism():
    """Every free node keeps a path to a support bond or a pinned node, cut after cut."""
    for pinned in (False, True):
        g = gen.piece(400, pinned=pinned, seed=21)
        broken = np.zeros(len(g.bonds), np.uint8)
        for r in range(30):
            cut = gen.safe_cuts(g, broken, 8, seed=r)
            assert len(set(cut.tolist())) == len(cut) and not broken[cut].any()
            broken[cut] = 1
        assert broken.sum() > len(g.bonds) // 4
        assert _grounded(g, broken)


def test_city_shapes_and_ground_storey_collapse():
    """City buildings: sizes in range and about the requested total; a ground-storey collapse cuts
    only bonds between the two lowest storeys under half the footprint and leaves no mechanism."""
    sh = gen.city_shapes(400, seed=21)
    n = np.array([a * b * c for a, b, c in sh])
    assert 4500 <= n.min() and n.max() <= 155000 and 10e6 < n.sum() < 14e6
    assert all(c >= a for a, b, c in sh)  # (taller than wide)
    nx, ny, nz = 6, 7, 9
    g = gen.building(nx, ny, nz, seed=3)
    cut = gen.ground_storey_cuts(g)
    assert len(cut) == (nx // 2) * ny
    z = g.nodes["c"][:, 2]
    zs = np.unique(z)
    for k in cut:
        a, b = int(g.bonds["a"][k]), int(g.bonds["b"][k])
        assert b >= 0 and sorted((z[a], z[b])) == [zs[0], zs[1]]
    broken = np.zeros(len(g.bonds), np.uint8)
    broken[cut] = 1
    assert _grounded(g, broken)


def test_hardware_queues_raised_before_the_library_loads():
    """A server's engine and setup workers each run on a stream of their own: the package asks HIP for
    at least 8 hardware queues (read when the runtime starts) before it loads the library."""
    from opus_destruct_2_amd import stress

    stress.lib()
    assert int(os.environ["GPU_MAX_HW_QUEUES"]) >= 8
