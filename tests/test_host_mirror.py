"""The C++ host mirror (include/rest_tensors.hpp): RIFull / MatrixFull /
MatrixUpper / _dgemm / davidson_solve with the reference's names and
semantics, exercised by tests/cpp/test_host_mirror.cpp (built by
rest_tensors_amd.build next to the library objects)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "rest_tensors_amd", "_build", "test_host_mirror")


def run(*args):
    assert os.path.exists(BIN), f"{BIN} is not built (run rest_tensors_amd.build)"
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    return r.stdout


def test_host_mirror_cpu_checks():
    out = run("--cpu-only")
    assert "ok rifull_from_vec_short_panics" in out


@pytest.mark.gpu
def test_host_mirror_on_gpu():
    out = run()
    for name in ("dgemm_doc_nn", "dgemm_doc_tn", "dgemm_doc_nt", "ao2mo_bench_closed_form",
                 "to_matrixupper_doc_order", "davidson_residual"):
        assert f"ok {name}" in out
