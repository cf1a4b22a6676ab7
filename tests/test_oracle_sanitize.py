"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
`sanitize`, driver oracle/san_driver.c): every batched entry point at small sizes, with the toy
PDFs and with a grid file, must run clean (a sanitizer report aborts with a non-zero status)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pdfgrid_util as PG  # noqa: E402


def test_oracle_runs_clean_under_asan_ubsan(tmp_path):
    oracle = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-C", oracle, "sanitize"], check=True, capture_output=True)
    path = str(tmp_path / "grid.dat")
    PG.write_lhagrid1(path, PG.nnpdf_like_x_grid(20, 20, 1e-7), [PG.q_grid(1.0, 100.0, 10), [100.0, 1e3, 1e4]],
                      PG.cubic_fn(5))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    for args in ([], [path]):
        r = subprocess.run([os.path.join(oracle, "_build", "oracle_san")] + args, env=env, capture_output=True,
                           text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-3000:]
        assert "checksum" in r.stdout
