"""The wave-parallel quicksort's partition rule, on the CPU.

sort_desc_wave (engine_gpu.hip) replaces the serial scan of the reference's
quicksort (runner.c:110-174) by ranks: the k-th swap pairs the k-th element
from the left with key <= pivot and the k-th from the right with key >= pivot
while the former lies at or before the latter, and the scan's stopping
points follow from the next pair.  tests/csrc/sort_partition_check.c writes
that rule as plain loops and compares the final order with the serial
quicksort on random lists, many with repeated keys.  (The GPU kernel itself
is compared bit for bit with the oracle's sort lists in test_gpu_parity.)
"""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "csrc" / "sort_partition_check.c"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_rank_partition_equals_serial_quicksort(tmp_path):
    exe = tmp_path / "sort_partition_check"
    subprocess.run(["gcc", "-O2", "-o", str(exe), str(SRC)], check=True)
    r = subprocess.run([str(exe), "200000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok 200000"), r.stdout
