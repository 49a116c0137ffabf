"""The batch engine's device-memory arena (csrc/arena.hpp), checked on the
host: randomised allocation/release against a byte map (no overlap,
alignment, containment, first fit finds any free run, full coalescing)."""
import os
import subprocess

import pytest

from conftest import PKG, REPO


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_arena_randomised(tmp_path, seed):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(PKG, "csrc"),
                    os.path.join(REPO, "tests", "native", "arena_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(seed)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
