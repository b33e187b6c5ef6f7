"""C ABI of libwlc_mde.so on the host (no GPU call): the library loads, exports
every function include/wlc_mde.h declares, and rejects bad parameters before it
touches a device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "wlc_mde.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wlc_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol(wlc):
    lib = wlc.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(wlc.EXPORTS)


def _params(wlc, **kw):
    p = wlc.WlcParams()
    base = dict(nx=8, ny=8, nz=8, box=(2.0, 2.0, 2.0), nblk=3, Ns=12, Nblk=(4, 4, 4), Lmax=(2, 2, 2),
                lam=(1.0, 1.0, 1.0))
    base.update(kw)
    p.nx, p.ny, p.nz = base["nx"], base["ny"], base["nz"]
    for a in range(3):
        p.box[a] = base["box"][a]
    p.nblk, p.Ns = base["nblk"], base["Ns"]
    for b in range(3):
        p.Nblk[b] = base["Nblk"][b]
        p.Lmax[b] = base["Lmax"][b]
        p.lam[b] = base["lam"][b]
    return p


@pytest.mark.parametrize("bad", [
    dict(nx=7), dict(ny=10), dict(nz=0), dict(nx=512), dict(box=(2.0, -1.0, 2.0)), dict(nblk=4),
    dict(nblk=0), dict(Ns=13), dict(Nblk=(4, 0, 8)), dict(Lmax=(2, 9, 2)), dict(Lmax=(-1, 2, 2)),
    dict(lam=(1.0, float("nan"), 1.0)), dict(Ns=2, Nblk=(1, 1, 0), nblk=2),
    # half planes beyond the 160 KB of LDS a plane kernel holds
    dict(nx=144, ny=144), dict(nx=192, ny=128), dict(nx=256, ny=256),
])
def test_create_rejects_bad_parameters(wlc, bad):
    lib = wlc.load()
    p = _params(wlc, **bad)
    h = ctypes.c_void_p()
    rc = lib.wlc_mde_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1  # WLC_EINVAL
    assert not h.value
    assert lib.wlc_mde_last_error()


@pytest.mark.parametrize("nx,ny", [(144, 144), (192, 128), (256, 256)])
def test_oversized_planes_name_the_limit(wlc, nx, ny):
    lib = wlc.load()
    p = _params(wlc, nx=nx, ny=ny)
    h = ctypes.c_void_p()
    assert lib.wlc_mde_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"LDS" in lib.wlc_mde_last_error()


def test_reserved_fields_must_be_zero(wlc):
    lib = wlc.load()
    p = _params(wlc)
    p.reserved[3] = 1
    h = ctypes.c_void_p()
    assert lib.wlc_mde_create(ctypes.byref(p), ctypes.byref(h)) == -1


@pytest.mark.parametrize("k", [-1, 1, 2])
def test_checkpoint_interval_is_validated(wlc, k):
    lib = wlc.load()
    p = _params(wlc)
    p.ckpt_every = k
    h = ctypes.c_void_p()
    assert lib.wlc_mde_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"ckpt_every" in lib.wlc_mde_last_error()


@pytest.mark.parametrize("v", [-1, 2])
def test_spectrum_precision_flag_is_validated(wlc, v):
    lib = wlc.load()
    p = _params(wlc)
    p.spectrum_fp32 = v
    h = ctypes.c_void_p()
    assert lib.wlc_mde_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"spectrum_fp32" in lib.wlc_mde_last_error()


def test_params_layout_matches_the_header(wlc):
    # int nx,ny,nz; double box[3]; int nblk,Ns,Nblk[3],Lmax[3]; double lambda[3];
    # int keep_qstar, ckpt_every, spectrum_fp32, reserved[5]
    assert ctypes.sizeof(wlc.WlcParams) == 4 * 3 + 4 + 8 * 3 + 4 * 8 + 8 * 3 + 4 * 8
    assert wlc.WlcParams.ckpt_every.offset == wlc.WlcParams.keep_qstar.offset + 4
    assert wlc.WlcParams.spectrum_fp32.offset == wlc.WlcParams.ckpt_every.offset + 4
    assert wlc.WlcParams.reserved.offset == wlc.WlcParams.spectrum_fp32.offset + 4


def test_null_arguments_are_errors(wlc):
    lib = wlc.load()
    h = ctypes.c_void_p()
    assert lib.wlc_mde_create(None, ctypes.byref(h)) == -1
    assert lib.wlc_mde_forward(None) == -1
    assert lib.wlc_mde_backward(None) == -1
    assert lib.wlc_mde_set_w(None, None) == -1
    assert lib.wlc_mde_sync(None) == -1
    assert lib.wlc_mde_slice_M(None, 0) == -1
    lib.wlc_mde_destroy(None)


def test_block_lengths_follow_reference_rounding(wlc):
    # NA = NINT(fA Ns), NB = NINT(fB Ns), NC = rest (initialize_mod.f90:141-145)
    assert wlc.block_lengths(100, [1 / 3] * 3) == [33, 33, 34]
    assert wlc.block_lengths(200, [1 / 3] * 3) == [67, 67, 66]
    assert wlc.block_lengths(256, [1 / 3] * 3) == [85, 85, 86]
    assert wlc.block_lengths(100, [0.25, 0.75]) == [25, 75]


def test_compiled_tables_match_host_setup():
    """The build regenerates wlc_tables.inc from the host set-up code; the committed
    copy must be current."""
    import subprocess
    import tempfile
    csrc = os.path.join(ROOT, "wormlike-chain-triblock_amd", "csrc")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(csrc, "gen_tables.cpp"),
                               os.path.join(csrc, "sh_setup.cpp")])
        out = os.path.join(d, "t.inc")
        subprocess.check_call([exe, out])
        assert open(out).read() == open(os.path.join(csrc, "wlc_tables.inc")).read()
