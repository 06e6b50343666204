"""Host checks of the sparse path's plan (zero-rk_amd/csrc/sparse_plan.cpp),
built with g++ from the package sources and tests/native/sparse_plan_check.cpp:
index ranges, level orders (every row after the rows it depends on), the
trailing block, and the GPU kernel's factorisation and solves emulated in
FP64 on a random matrix with the plan's pattern (residual against the matrix
itself; block entries outside the filled pattern stay exactly zero)."""
import os
import subprocess

import pytest

from conftest import ROOT, mech_paths

CSRC = os.path.join(ROOT, "zero-rk_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("plan") / "sparse_plan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(ROOT, "tests", "native", "sparse_plan_check.cpp"),
                    os.path.join(CSRC, "sparse_plan.cpp"), os.path.join(CSRC, "ck_mech.cpp"),
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("name,emulate", [("synth_10sp", 1), ("synth_53sp", 1),
                                          ("synth_160sp", 1), ("ref_h2", 1), ("ref_dme", 1),
                                          ("synth_2000sp", 1)])
def test_sparse_plan(checker, name, emulate):
    r = subprocess.run([checker, *mech_paths(name), str(emulate)], capture_output=True, text=True,
                       timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " ok" in r.stdout
