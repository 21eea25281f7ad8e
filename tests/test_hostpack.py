"""The host loops that move particles between the cells and the device's
cell-ordered arrays (space_setcells, space_pack, space_add_forces), run in
parallel over cells with OpenMP: tests/csrc/hostpack_check.c builds them
from the library's sources and checks every field, the id maps, the packed
arrays and the force write-back after a random new layout, with 1, 3 and 8
threads."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "mdcore-history_amd" / "src"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_host_packing_loops(tmp_path):
    exe = tmp_path / "hostpack_check"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fopenmp", f"-I{ROOT / 'include'}", f"-I{SRC}", "-o", str(exe),
                    str(ROOT / "tests" / "csrc" / "hostpack_check.c"), str(SRC / "space.c"), str(SRC / "hostpack.c"),
                    str(SRC / "errs.c"), "-lm"], check=True)
    for threads in (1, 3, 8):
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        assert r.returncode == 0 and r.stdout.strip() == "ok", (threads, r.stdout, r.stderr)
