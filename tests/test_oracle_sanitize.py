"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/sanitize_check.c runs every oracle entry point (both precisions, every boundary
mode, one and two OpenMP threads) on small inputs; built here with
-fsanitize=address,undefined it must run to completion with no report.
"""
import os
import pathlib
import shutil
import subprocess

import pytest

ORACLE = pathlib.Path(__file__).resolve().parent.parent / "oracle"
FLAGS = ["-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-D_GNU_SOURCE", "-Wall", "-Wextra", "-Werror"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    src = ORACLE / "hydro_oracle.c"
    steps = [
        ["gcc", *FLAGS, "-c", str(src), "-o", str(tmp_path / "f64.o")],
        ["gcc", *FLAGS, "-DORACLE_F32", "-c", str(src), "-o", str(tmp_path / "f32.o")],
        ["gcc", *FLAGS, str(ORACLE / "sanitize_check.c"), str(tmp_path / "f64.o"), str(tmp_path / "f32.o"), "-o",
         str(tmp_path / "check"), "-lm"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stderr}"
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(tmp_path / "check"), str(tmp_path / "scratch.csv")], capture_output=True, text=True,
                       env=env, timeout=300)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report
    assert "Sanitizer" not in report and "runtime error" not in report, report
    assert not (tmp_path / "scratch.csv").exists()
