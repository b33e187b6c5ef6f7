"""Field files 'W<rank>.dat' of the Fortran driver (module field_io): one line per grid
point, K then the block fields (the reference's W_dump, SCFT_dump.f90:273-293, and readW,
init_array_2dcomp_mod.f90:1070-1091).  Checked without a GPU through field_io_check."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "wormlike-chain-triblock_amd", "fortran", "field_io_check")

pytestmark = pytest.mark.skipif(not os.path.exists(CHECK),
                                reason="Fortran field_io_check not built (no flang)")


def run(n, nblk, inp, wdat, out):
    return subprocess.run([CHECK, str(n), str(nblk), inp, str(wdat), str(out)],
                          capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("nblk", [2, 3])
def test_dump_and_read_back_are_exact(tmp_path, nblk):
    """Every double, including signed zeros, subnormals and the extremes, survives the
    text round trip bit for bit; an independent reader (numpy.loadtxt) sees the same values
    and the indices 0..N-1 in the first column."""
    n = 4096
    rng = np.random.default_rng(7)
    w = rng.standard_normal((nblk, n)) * 10.0 ** rng.integers(-300, 300, (nblk, n))
    w[0, :6] = [0.0, -0.0, 5e-324, 1.7976931348623157e308, -2.2250738585072014e-308, 1.0 / 3.0]
    w.tofile(tmp_path / "in.bin")
    r = run(n, nblk, str(tmp_path / "in.bin"), tmp_path / "W0.dat", tmp_path / "out.bin")
    assert r.returncode == 0, r.stdout + r.stderr
    back = np.fromfile(tmp_path / "out.bin").reshape(nblk, n)
    assert np.array_equal(back.view(np.uint64), w.view(np.uint64))
    t = np.loadtxt(tmp_path / "W0.dat")
    assert t.shape == (n, nblk + 1)
    assert np.array_equal(t[:, 0], np.arange(n))
    assert np.array_equal(np.ascontiguousarray(t[:, 1:].T).view(np.uint64), w.view(np.uint64))


def test_reads_list_directed_files_written_elsewhere(tmp_path):
    """The reader is list-directed like the reference's: free spacing, exponent letters D
    and E, fewer digits, commas as separators."""
    n, nblk = 5, 2
    vals = np.array([[0.5, -1.25], [3.0e-3, 7.0], [-0.0, 1.0e10], [2.5, -2.5], [1.0, 2.0]])
    lines = ["           0  0.50000000000000000      -1.2500000000000000",
             "1 3.0D-3 7",
             "  2, -0.0, 1.0E+10",
             "3\t2.5\t-2.5",
             "   4   1.0000000000000000        2.0000000000000000     "]
    (tmp_path / "W0.dat").write_text("\n".join(lines) + "\n")
    r = run(n, nblk, "-", tmp_path / "W0.dat", tmp_path / "out.bin")
    assert r.returncode == 0, r.stdout + r.stderr
    back = np.fromfile(tmp_path / "out.bin").reshape(nblk, n)
    assert np.array_equal(back, vals.T)


def test_index_out_of_order_is_an_error(tmp_path):
    (tmp_path / "W0.dat").write_text("0 1.0 2.0\n2 3.0 4.0\n1 5.0 6.0\n")
    r = run(3, 2, "-", tmp_path / "W0.dat", tmp_path / "out.bin")
    assert r.returncode != 0 and "status -1" in r.stdout


def test_short_file_is_an_error(tmp_path):
    (tmp_path / "W0.dat").write_text("0 1.0 2.0\n1 3.0 4.0\n")
    r = run(3, 2, "-", tmp_path / "W0.dat", tmp_path / "out.bin")
    assert r.returncode != 0 and "read_w_file failed" in r.stdout
