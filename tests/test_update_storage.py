"""The lattice factor's shared update storage, checked by brute force on the CPU.

A supernode's update matrix is written when its level of the multifrontal factor runs and read
when its parent's level runs; the plan gives updates whose lives do not meet the same storage.
tests/native/update_storage_check.cpp (compiled here with g++ against the library's own
cells_amd/csrc/symbolic.cpp) builds the supernodes of lattice worlds of the batched sizes (the
relaxed plan the device factors), shares the storage and tests every pair of updates alive at a
common level for overlap, and every update for staying inside the storage. The storage must also
be smaller than the updates laid end to end.
"""

import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from cells_amd import symbolic
from cells_amd.world import lattice

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    out = tmp_path_factory.mktemp("native") / "update_storage_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "cells_amd" / "csrc"), "-o", str(out),
                    str(ROOT / "tests" / "native" / "update_storage_check.cpp"),
                    str(ROOT / "cells_amd" / "csrc" / "symbolic.cpp")], check=True)
    return out


def _relaxed_pattern(n):
    snd, rcv = lattice(n)
    rows = np.concatenate([np.arange(n), np.maximum(snd, rcv)])
    cols = np.concatenate([np.arange(n), np.minimum(snd, rcv)])
    order, pr, pc = symbolic.relaxed_padding(n, rows, cols, 128, 64)
    return (np.concatenate([rows, np.maximum(pr, pc)]), np.concatenate([cols, np.minimum(pr, pc)]), order)


@pytest.mark.parametrize("n", [1, 2, 7, 256, 777, 2000, 3001, 4096])
def test_live_update_matrices_never_share_storage(checker, tmp_path, n):
    rows, cols, order = _relaxed_pattern(n)
    path = tmp_path / f"pattern_{n}.txt"
    with open(path, "w") as f:
        f.write(f"{n} {rows.size}\n")
        for a in (rows, cols, order):
            f.write(" ".join(map(str, a.tolist())) + "\n")
    line = subprocess.run([str(checker), str(path)], check=True, capture_output=True, text=True).stdout.split()
    got = dict(zip(line[0::2], map(int, line[1::2])))
    assert got["n"] == n
    assert got["overlaps"] == 0
    assert got["peak"] <= got["total"]
    if n >= 256:
        assert got["peak"] * 2 < got["total"]
