"""The committed kernel-B variant table (csrc/wlc_tune.inc, written by autotune.py).

The library compiles the table in and takes its row for (nz, Lmax, spectrum precision) as the
z-solve variant; a row the library cannot launch for its shape would only show on the device.
These checks keep the table consistent with the autotuner that writes it: one row per tuned
shape, every chosen variant among the candidates the autotuner times for that shape, and the
writer's line format read back unchanged.
"""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wormlike-chain-triblock_amd")
sys.path.insert(0, PKG)

import autotune  # noqa: E402

ROW = re.compile(r"^\s*\{\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*([0-9.]+)\},\s*$")


def read_rows(text):
    rows = []
    for line in text.splitlines():
        if not line.strip() or line.lstrip().startswith("//"):
            continue
        m = ROW.match(line)
        assert m, f"unparsable table line: {line!r}"
        rows.append(tuple(int(x) for x in m.groups()[:7]) + (float(m.group(8)),))
    return rows


@pytest.fixture(scope="module")
def rows():
    with open(os.path.join(PKG, "csrc", "wlc_tune.inc")) as f:
        return read_rows(f.read())


def test_one_row_per_tuned_shape(rows):
    keys = [r[:3] for r in rows]
    assert len(keys) == len(set(keys)), "duplicate (nz, Lmax, fp32) rows"
    space = {(nz, L, sp) for nz, Ls, sps in autotune.SPACE for L in Ls for sp in sps}
    assert set(keys) == space


def test_every_choice_is_a_timed_candidate(rows):
    for nz, L, sp, gen, split, occ, tx, us in rows:
        assert (gen, split, occ, tx) in autotune.candidates(nz, L), (nz, L, sp)
        assert us > 0.0


def test_candidates_respect_the_library_limits():
    # the library accepts split 0..3, occupancy 1 or 2, and tx with 1 <= tx and tx nz <= 256
    # for the generic kernel; the specialised kernels exist for the shapes listed here
    special_split = {64: range(5, 8), 24: range(4, 8), 32: range(4, 8), 48: range(4, 8),
                     128: range(5, 9)}
    for nz, Ls, _ in autotune.SPACE:
        for L in Ls:
            for gen, split, occ, tx in autotune.candidates(nz, L):
                assert gen in (0, 1) and occ in (1, 2) and 0 <= split <= 3
                if gen:
                    assert 1 <= tx and tx * nz <= 256
                    if split:
                        assert 4 <= L <= 7
                else:
                    assert nz in (24, 32, 48, 64, 128)
                    if split:
                        assert L in special_split[nz], (nz, L, split)
                    if split == 3:
                        assert nz in (24, 32, 48, 64)
                    if split == 2:
                        assert nz == 64


def test_writer_format_reads_back():
    row = (48, 6, 0, 0, 3, 1, 1, 54.0)
    line = "    {%d, %d, %d, %d, %d, %d, %d, %.1f},\n" % row
    assert read_rows("// header\n" + line) == [row]
