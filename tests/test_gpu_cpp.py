"""evolve_flavor through the C++ host layer (emu_amd/cpp, main.cpp:37-227 on
this build), called in-process through libemu_cpp.so: the plain per-function
loop and the fused FP64 loop against the FP64 oracle at 1e-12, the fused
FP32-bulk loop within 2e-5, and the Renormalize abort as an error status."""
import ctypes
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cpp():
    import torch
    assert torch.cuda.is_available()
    from emu_amd import capi
    capi.lib()
    lib = ctypes.CDLL(os.path.join(ROOT, "emu_amd", "libemu_cpp.so"))
    lib.emu_cpp_evolve_flavor.restype = ctypes.c_int
    lib.emu_cpp_evolve_flavor.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                          ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_int]
    return lib


def run(lib, deck_path, path, dump):
    out = (ctypes.c_double * 6)()
    err = ctypes.create_string_buffer(1024)
    arr = (ctypes.c_char_p * 1)()
    code = lib.emu_cpp_evolve_flavor(str(deck_path).encode(), arr, 0, path.encode(),
                                     str(dump).encode() if dump else None, 1, out, err, 1024)
    return code, list(out), err.value.decode()


def read_dump(path, nf):
    with open(path, "rb") as f:
        assert f.read(8) == b"EMUDUMP1"
        hdr_nf, na = np.frombuffer(f.read(8), dtype=np.int32)
        np_, = np.frombuffer(f.read(8), dtype=np.int64)
        time, = np.frombuffer(f.read(8), dtype=np.float64)
        ids = np.frombuffer(f.read(8 * np_), dtype=np.int64)
        pos = np.frombuffer(f.read(24 * np_), dtype=np.float64).reshape(3, np_)
        rdata = np.frombuffer(f.read(8 * na * np_), dtype=np.float64).reshape(na, np_)
    assert hdr_nf == nf
    return ids, pos.T, rdata.T, time


@pytest.mark.parametrize("path,tol", [("plain", 1e-12), ("fused-fp64", 1e-12), ("fused-fp32", 2e-5)])
@pytest.mark.parametrize("nf", [2, 3])
def test_cpp_evolve_flavor_matches_oracle(cpp, oracle, tmp_path, nf, path, tol):
    nsteps = 3
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4, nsteps=nsteps, end_time=1.0)
    deck_path = tmp_path / "inputs"
    deck_path.write_text(cases.deck_text(d))
    dump = tmp_path / "final.bin"
    code, out, err = run(cpp, deck_path, path, dump)
    assert code == 0, err
    assert int(out[0]) == nsteps
    # the oracle: same initial conditions, same loop
    deck = cases.oracle_deck(oracle, d)
    prm = deck.params()
    parts = oracle.init_particles(deck)
    g = oracle.Grid(nf, prm)
    vv = oracle.vvac_max(nf, prm, parts)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    e, dt = oracle.compute_dt(nf, prm, g, vv)
    t = 0.0
    for _ in range(nsteps):
        e, dt_next, bad = oracle.rk4_step(nf, prm, parts, g, dt, vv)
        assert e == 0
        t += dt
        dt = dt_next
    ids, pos, rdata, time = read_dump(dump, nf)
    assert int(out[5]) == len(parts)
    want = parts[np.argsort(parts["id"])]
    np.testing.assert_array_equal(ids, want["id"])
    assert time == pytest.approx(t, rel=1e-13)
    # FP32-bulk keeps positions in FP64 but the momenta in FP32, so the
    # velocity carries FP32 rounding
    np.testing.assert_allclose(pos, want["pos"], rtol=1e-13 if tol < 1e-10 else 1e-7, atol=0)
    got = want.copy()
    got["rdata"] = rdata
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < tol


def test_cpp_renormalize_abort_is_a_status(cpp, tmp_path):
    """The reference aborts when Renormalize's corrections exceed 100
    maxError; through the C entry point that is a status code and a message."""
    d = cases.fast_flavor_dict(nf=2, ncell=(2, 2, 2), nphi=4, nsteps=2, st4_amplitude=1e-2, maxError=1e-300)
    deck_path = tmp_path / "inputs"
    deck_path.write_text(cases.deck_text(d))
    code, out, err = run(cpp, deck_path, "plain", None)
    assert code in (1, 2, 3) and "Renormalize" in err


def test_cpp_ssprk3_plain_and_fused_refusal(cpp, oracle, tmp_path):
    """integration.rk.type = 3 through the C++ TimeIntegrator against the
    oracle's SSPRK3 step; the fused path (RK4 only) refuses the deck."""
    nf, nsteps = 2, 3
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4, nsteps=nsteps, end_time=1.0)
    d["integration.type"], d["integration.rk.type"] = 1, 3
    deck_path = tmp_path / "inputs"
    deck_path.write_text(cases.deck_text(d))
    dump = tmp_path / "final.bin"
    code, out, err = run(cpp, deck_path, "plain", dump)
    assert code == 0, err
    deck = cases.oracle_deck(oracle, d)
    prm = deck.params()
    parts = oracle.init_particles(deck)
    g = oracle.Grid(nf, prm)
    vv = oracle.vvac_max(nf, prm, parts)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    e, dt = oracle.compute_dt(nf, prm, g, vv)
    for _ in range(nsteps):
        e, dt, bad = oracle.rk_step(nf, prm, parts, g, dt, vv, 3)
        assert e == 0
    ids, pos, rdata, time = read_dump(dump, nf)
    want = parts[np.argsort(parts["id"])]
    np.testing.assert_allclose(pos, want["pos"], rtol=1e-13, atol=0)
    assert cases.max_rel_err(rdata, want["rdata"]) < 1e-12
    code, out, err = run(cpp, deck_path, "fused-fp64", None)
    assert code == 12 and "RK4" in err
