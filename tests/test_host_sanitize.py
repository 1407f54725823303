"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

* The binary Gmsh 4.1 reader (arcanefem_amd/csrc/femutils/Mesh.hpp) parses
  every mesh fixture, and corrupted copies of each (truncations, byte flips,
  overwritten 64-bit counts) either parse or raise FatalError -- no crash,
  no out-of-bounds access, no oversized allocation, no undefined behaviour.
* The CPU oracle (oracle/afem_oracle.c) runs its pattern, assembly, RHS,
  transformations, toCsr, value lookup, SpMV and Jacobi-PCG clean.

The harnesses are built with g++ / gcc (host code only, no device code).
"""

import glob
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

HERE = os.path.join(ROOT, "tests", "host_sanitize")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not available")


def _run(cmd, env=None, timeout=300):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    return p.stdout


@pytest.fixture(scope="module")
def gmsh_fuzz(tmp_path_factory):
    _need("g++")
    out = str(tmp_path_factory.mktemp("san") / "gmsh_fuzz")
    _run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), "-I",
          os.path.join(ROOT, "arcanefem_amd", "csrc"), os.path.join(HERE, "gmsh_fuzz.cpp"), "-o", out])
    return out


def test_gmsh_reader_parses_every_fixture_clean(gmsh_fuzz):
    files = sorted(glob.glob(os.path.join(GOLDEN, "*.msh")))
    assert len(files) >= 10
    out = _run([gmsh_fuzz, "parse", *files])
    assert out.count("nodes=") == len(files)


def test_gmsh_reader_rejects_corrupt_files_without_crashing(gmsh_fuzz, tmp_path):
    env = dict(os.environ, HARNESS_TMP=str(tmp_path))
    for k, f in enumerate(sorted(glob.glob(os.path.join(GOLDEN, "*.msh")))):
        out = _run([gmsh_fuzz, "mutate", str(1000 + k), "80", f], env=env)
        parsed = int(out.split("parsed=")[1].split()[0])
        rejected = int(out.split("rejected=")[1].split()[0])
        assert parsed + rejected == 80 and rejected > 0, out


def test_oracle_runs_clean_under_sanitizers(tmp_path):
    _need("gcc")
    out = str(tmp_path / "oracle_asan")
    _run(["gcc", "-std=c11", *SAN, "-fopenmp", "-I", os.path.join(ROOT, "oracle"),
          os.path.join(HERE, "oracle_asan.c"), os.path.join(ROOT, "oracle", "afem_oracle.c"), "-o", out, "-lm"])
    res = _run([out])
    assert res.count(" ok") == 2
