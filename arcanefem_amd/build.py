"""Builds libafem.so (HIP kernels + C ABI + C++ host layer) for gfx950.

Usage: python -m arcanefem_amd.build [--jobs N] [--verbose]

Every HIP source is compiled by hipcc for gfx950 only, with packed FP32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops),
into _build/, then linked in-tree as arcanefem_amd/libafem.so.  The C++
programs under arcanefem_amd/csrc/apps/ are linked against it.
"""

import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libafem.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _sources():
    out = []
    for d in (CSRC, os.path.join(CSRC, "femutils")):
        if not os.path.isdir(d):
            continue
        for f in sorted(os.listdir(d)):
            if f.endswith(".hip") or f.endswith(".cpp"):
                out.append(os.path.join(d, f))
    return out


def _apps():
    d = os.path.join(CSRC, "apps")
    if not os.path.isdir(d):
        return []
    return [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".cpp")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(jobs=None, verbose=False):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    objs = []
    cmds = []
    for s in srcs:
        rel = os.path.relpath(s, CSRC).replace(os.sep, "_")
        o = os.path.join(BUILD, rel + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmds.append([HIPCC] + COMMON + lang + ["-c", s, "-o", o])
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(lambda c: _run(c, verbose), cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    for app in _apps():
        exe = os.path.join(HERE, "bin", os.path.splitext(os.path.basename(app))[0])
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        _run([HIPCC] + COMMON + [app, "-o", exe, "-L", HERE, "-lafem", "-Wl,-rpath,$ORIGIN/.."], verbose)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    try:
        print(build(a.jobs, a.verbose))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
