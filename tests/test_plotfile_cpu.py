"""Plotfiles (WritePlotFile / RecoverParticles, Source/IO.cpp:11-63) on the
CPU: a round trip through the C++ writer and both readers (emu_amd.plotfile
and the C++ reader RecoverParticles uses), and the layout the reference's
readers expect (Scripts/visualization/amrex_plot_tools.py:12-88 key lists and
its particle reader's header fields)."""
import ctypes
import os

import numpy as np
import pytest

from emu_amd import plotfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# the real-component keys of the reference's readers (amrex_plot_tools.py:12-88)
REF_KEYS = {
    2: ["pos_x", "pos_y", "pos_z", "time", "x", "y", "z", "pupx", "pupy", "pupz", "pupt", "N", "L", "f00_Re",
        "f01_Re", "f01_Im", "f11_Re", "Nbar", "Lbar", "f00_Rebar", "f01_Rebar", "f01_Imbar", "f11_Rebar"],
    3: ["pos_x", "pos_y", "pos_z", "time", "x", "y", "z", "pupx", "pupy", "pupz", "pupt", "N", "L", "f00_Re",
        "f01_Re", "f01_Im", "f02_Re", "f02_Im", "f11_Re", "f12_Re", "f12_Im", "f22_Re", "Nbar", "Lbar",
        "f00_Rebar", "f01_Rebar", "f01_Imbar", "f02_Rebar", "f02_Imbar", "f11_Rebar", "f12_Rebar", "f12_Imbar",
        "f22_Rebar"],
}


@pytest.fixture(scope="module")
def cpp():
    from emu_amd import capi
    capi.lib()
    lib = ctypes.CDLL(os.path.join(ROOT, "emu_amd", "libemu_cpp.so"))
    P = ctypes.c_void_p
    lib.emu_cpp_write_plotfile.argtypes = [ctypes.c_char_p, ctypes.c_int, P, P, P, ctypes.c_double, ctypes.c_int, P,
                                           ctypes.c_int64, P, P, P, ctypes.c_char_p, ctypes.c_int]
    lib.emu_cpp_read_plotfile.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_int64, P, P, P, ctypes.c_char_p, ctypes.c_int]
    return lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def write(lib, d, nf, ncell, time, step, grid, ids, pos, rdata):
    err = ctypes.create_string_buffer(512)
    n = np.array(ncell, dtype=np.int32)
    plo = np.zeros(3)
    phi = np.array([0.5 * c for c in ncell])
    code = lib.emu_cpp_write_plotfile(str(d).encode(), nf, ptr(n), ptr(plo), ptr(phi), time, step, ptr(grid),
                                      len(ids), ptr(ids), ptr(pos), ptr(rdata), err, 512)
    assert code == 0, err.value


@pytest.mark.parametrize("nf", [2, 3])
def test_plotfile_round_trip(cpp, tmp_path, nf):
    rng = np.random.default_rng(nf)
    ncell = (3, 4, 5)
    ncomp = 3 + 8 * nf * nf
    na = 12 + 2 * nf * nf
    grid = rng.standard_normal((ncomp, ncell[2], ncell[1], ncell[0]))
    n = 57
    ids = rng.permutation(np.arange(1, n + 1)).astype(np.int64)
    pos = rng.uniform(0, 1.5, (3, n))
    rdata = rng.standard_normal((na, n))
    d = tmp_path / "plt00007"
    write(cpp, d, nf, ncell, 3.25e-11, 7, grid, ids, pos, rdata)
    h = plotfile.read_header(str(d))
    assert h["time"] == 3.25e-11 and h["step"] == 7 and h["ncell"] == list(ncell) and h["dim"] == 3
    assert h["dx"] == [0.5, 0.5, 0.5]
    names, g = plotfile.read_grid(str(d))
    assert names[:3] == ["rho", "T", "Ye"] and len(names) == ncomp
    np.testing.assert_array_equal(g, grid)
    rid, rr, rnames = plotfile.read_particles(str(d))
    assert rnames == REF_KEYS[nf]
    order = np.argsort(ids)
    np.testing.assert_array_equal(rid, ids[order])          # written in id order
    np.testing.assert_array_equal(rr[:, :3], pos.T[order])
    np.testing.assert_array_equal(rr[:, 3:], rdata.T[order])
    # the C++ reader of RecoverParticles
    t, s, cnt = ctypes.c_double(), ctypes.c_int(), ctypes.c_int64()
    err = ctypes.create_string_buffer(512)
    assert cpp.emu_cpp_read_plotfile(str(d).encode(), nf, ctypes.byref(t), ctypes.byref(s), ctypes.byref(cnt), 0,
                                     None, None, None, err, 512) == 0
    assert (t.value, s.value, cnt.value) == (3.25e-11, 7, n)
    bi, bp, br = np.zeros(n, dtype=np.int64), np.zeros((3, n)), np.zeros((na, n))
    assert cpp.emu_cpp_read_plotfile(str(d).encode(), nf, ctypes.byref(t), ctypes.byref(s), ctypes.byref(cnt), n,
                                     ptr(bi), ptr(bp), ptr(br), err, 512) == 0
    np.testing.assert_array_equal(bi, ids[order])
    np.testing.assert_array_equal(bp, pos[:, order])
    np.testing.assert_array_equal(br, rdata[:, order])


def test_plotfile_layout_matches_the_reference_readers(cpp, tmp_path):
    """The fields amrex_plot_tools.AMReXParticleHeader parses, in its order,
    and an IEEE little-endian FAB for the grid."""
    nf = 2
    grid = np.zeros((35, 1, 1, 2))
    ids = np.array([2, 1], dtype=np.int64)
    pos = np.zeros((3, 2))
    rdata = np.zeros((20, 2))
    d = tmp_path / "plt00000"
    write(cpp, d, nf, (2, 1, 1), 0.0, 0, grid, ids, pos, rdata)
    lines = open(d / "neutrinos" / "Header").read().split("\n")
    assert lines[0] == "Version_Two_Dot_Zero_double"
    assert int(lines[1]) == 3 and int(lines[2]) == 20
    assert lines[3:23] == REF_KEYS[2][3:]
    assert int(lines[23]) == 0            # no extra int components
    assert int(lines[24]) == 1            # checkpoint: (id, cpu) ints are present
    assert int(lines[25]) == 2            # particles
    assert int(lines[27]) == 0 and int(lines[28]) == 1  # finest level, grids on level 0
    assert [int(v) for v in lines[29].split()] == [0, 2, 0]
    raw = open(d / "neutrinos" / "Level_0" / "DATA_00000", "rb").read()
    assert len(raw) == 2 * 2 * 4 + 2 * 23 * 8
    assert list(np.frombuffer(raw[:16], dtype="<i4")) == [1, 0, 2, 0]
    fab = open(d / "Level_0" / "Cell_D_00000", "rb").readline().decode()
    assert fab.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))((0,0,0) (1,0,0) (0,0,0)) 35")
    assert os.path.exists(d / "Level_0" / "Cell_H") and os.path.exists(d / "neutrinos" / "Level_0" / "Particle_H")


def test_plotfile_reader_errors(cpp, tmp_path):
    err = ctypes.create_string_buffer(512)
    t, s, cnt = ctypes.c_double(), ctypes.c_int(), ctypes.c_int64()
    code = cpp.emu_cpp_read_plotfile(str(tmp_path / "missing").encode(), 2, ctypes.byref(t), ctypes.byref(s),
                                     ctypes.byref(cnt), 0, None, None, None, err, 512)
    assert code != 0 and b"Header" in err.value
    # a 2-flavour plotfile read as 3 flavours
    d = tmp_path / "plt00001"
    write(cpp, d, 2, (1, 1, 2), 0.0, 1, np.zeros((35, 2, 1, 1)), np.array([1], dtype=np.int64), np.zeros((3, 1)),
          np.zeros((20, 1)))
    code = cpp.emu_cpp_read_plotfile(str(d).encode(), 3, ctypes.byref(t), ctypes.byref(s), ctypes.byref(cnt), 0,
                                     None, None, None, err, 512)
    assert code != 0 and b"NUM_FLAVORS" in err.value
