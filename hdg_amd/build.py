"""Builds libhdgx.so (gfx950) in-tree: one object per kernel configuration,
compiled in parallel with hipcc, plus the C-ABI layer.

    python -m hdg_amd.build            # full rebuild
    python -m hdg_amd.build --changed  # only objects older than their sources
"""
import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhdgx.so")
ARCH = os.environ.get("HDGX_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # packed FP32 instructions are kept out of every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def configs():
    txt = open(os.path.join(CSRC, "hdgx_launch.h")).read()
    body = txt[txt.index("#define HDGX_FOR_EACH_CFG"):txt.index("#define HDGX_FOR_EACH_FAMILY")]
    return [tuple(int(v) for v in m) for m in re.findall(r"X\((\d+),\s*(\d+),\s*(\d+)\)", body)]


def families():
    txt = open(os.path.join(CSRC, "hdgx_launch.h")).read()
    body = txt[txt.index("#define HDGX_FOR_EACH_FAMILY"):txt.index("namespace hdgx")]
    return [tuple(int(v) for v in m) for m in re.findall(r"X\((\d+),\s*(\d+)\)", body)]


def _sources():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(HERE, "..", "include", "hdgx.h")]


def _stale(obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(s) > t for s in _sources())


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(changed_only=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs_list = []
    for ph, nd, p in configs():
        obj = os.path.join(BUILD, f"inst_{ph}_{nd}_{p}.o")
        cmd = [HIPCC, *FLAGS, f"-DHDGX_PHYS={ph}", f"-DHDGX_ND={nd}", f"-DHDGX_P={p}", "-c",
               os.path.join(CSRC, "hdgx_inst.hip"), "-o", obj]
        jobs_list.append((obj, cmd))
    for ph, nd in families():
        obj = os.path.join(BUILD, f"mixed_{ph}_{nd}.o")
        cmd = [HIPCC, *FLAGS, f"-DHDGX_PHYS={ph}", f"-DHDGX_ND={nd}", "-c", os.path.join(CSRC, "mixed_inst.hip"),
               "-o", obj]
        jobs_list.append((obj, cmd))
    for name in ("hdgx_api", "trace", "mex_inplace", "mesh"):
        obj = os.path.join(BUILD, f"{name}.o")
        jobs_list.append((obj, [HIPCC, *FLAGS, "-c", os.path.join(CSRC, f"{name}.hip"), "-o", obj]))
    todo = [(o, c) for o, c in jobs_list if not changed_only or _stale(o)]
    n = jobs or min(len(todo) or 1, max(1, (os.cpu_count() or 2)), 16)
    with cf.ThreadPoolExecutor(n) as ex:
        for out in ex.map(lambda oc: _run(oc[1]), todo):
            if out.strip():
                print(out)
    objs = [o for o, _ in jobs_list]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--changed", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(a.changed, a.j))
