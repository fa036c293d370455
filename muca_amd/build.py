"""Build libmuca.so and the drop-in CLIs in-tree, for gfx950 only.

    python -m muca_amd.build          # or muca_amd.build.build()

Outputs (git-ignored, shipped with the tree):
    muca_amd/libmuca.so               the C-ABI library (include/muca.h)
    muca_amd/bin/multigpu_weight      drop-in for muca_delta_multigpu.cu's binary
    muca_amd/bin/produce_gpu          drop-in for produce2_multigpu.cu's binary
    muca_amd/libmuca_probe.so         VALU-ceiling microbenchmarks for bench.py
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BIN = os.path.join(PKG, "bin")
LIB = os.path.join(PKG, "libmuca.so")

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -ffp-contract=off: every float/double expression of the acceptance rounds
# in the order it is written (fma only where written explicitly).
# -packed-fp32-ops off: no v_pk_*_f32 in generated code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas"]

# largest first: the team kernel's instantiations dominate the build and are
# split into four units so that the units compile in parallel
LIB_SOURCES = ["muca_team_r7_f32.hip", "muca_team_r7_f64.hip", "muca_team_r10_f32.hip",
               "muca_team_r10_f64.hip", "muca_team.hip", "muca_lane.hip", "muca_kernels.hip",
               "muca_mlp.hip", "muca_batch.hip", "muca_api.cpp", "muca_noalloc.hip"]
CLI_SOURCES = {"multigpu_weight": "cli_weight.cpp", "produce_gpu": "cli_produce.cpp"}


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return p


def build_lib(verbose=False, jobs=None):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(os.path.join(PKG, "_build"), exist_ok=True)
    cmds, objs = [], []
    for src in LIB_SOURCES:
        obj = os.path.join(PKG, "_build", src + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd))
        cmds.append(cmd)
        objs.append(obj)
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose=False):
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name, src in CLI_SOURCES.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        exe = os.path.join(BIN, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), path,
               "-o", exe, "-L", PKG, "-lmuca", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        out.append(exe)
    return out


def build_probe(verbose=False):
    """libmuca_probe.so: microbenchmarks bench.py prices the update kernels
    against (not part of the libmuca ABI)."""
    out = os.path.join(PKG, "libmuca_probe.so")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "probe_valu.hip"), "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False):
    lib = build_lib(verbose)
    clis = build_cli(verbose)
    return [lib] + clis + [build_probe(verbose)]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
