"""Build recipe of the package's native libraries (in-tree, for gfx950):

  libsvxg.so    the device library: HIP kernels + host driver + C ABI (include/svxg.h)
  libsvxgen.so  the deterministic synthetic-workload generator (CPU)
  tests/cpp/test_stress_facade  the reference's stress tests restated over include/svxg_stress.hpp
  tests/cpp/test_world_loops    World's loops over a structure's bonds against the same header

Kernels are compiled with -ffp-contract=off (assembly, P and the bond checks reproduce the
reference's operation order) and without packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SVXG_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
SOURCES = ["solve_kernels.hip", "setup_kernels.hip", "judge_kernels.hip", "section_kernels.hip", "lattice_kernels.hip",
           "extract_kernels.hip", "host.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_device(jobs=3):
    os.makedirs(BUILD, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        procs.append(subprocess.Popen([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    failed = False
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
        if out and ("error" in out):
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc failed")
    lib = os.path.join(HERE, "libsvxg.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib + ".tmp"] + objs)
    os.replace(lib + ".tmp", lib)
    return lib


def build_generator():
    lib = os.path.join(HERE, "libsvxgen.so")
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-o", lib + ".tmp", os.path.join(CSRC, "gen.cpp")])
    os.replace(lib + ".tmp", lib)
    return lib


def build_cpp_tests():
    """The C++ tests over include/svxg_stress.hpp: the reference's stress tests restated, and World's
    loops over a structure's bonds (linked with the workload generator for the block)."""
    outs = []
    for name, extra in (("test_stress_facade", []), ("test_world_loops", ["-lsvxgen"])):
        out = os.path.join(ROOT, "tests", "cpp", name)
        _run(["g++", "-std=c++20", "-O2", "-I", os.path.join(ROOT, "include"), "-o", out + ".tmp",
              os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-L", HERE, "-lsvxg"] + extra +
             [f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../opus_destruct_2_amd"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_all():
    build_generator()
    build_device()
    build_cpp_tests()


if __name__ == "__main__":
    build_all()
