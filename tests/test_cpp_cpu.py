"""The C++ host layer (emu_amd/cpp): its deck reader against the Python one
on every reference deck, its error behaviour, and the emu_main executable's
parameter dump.  No GPU work."""
import ctypes
import json
import os
import subprocess

import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBCPP = os.path.join(ROOT, "emu_amd", "libemu_cpp.so")
MAIN = os.path.join(ROOT, "emu_amd", "bin", "emu_main")


@pytest.fixture(scope="module")
def cpp():
    if not os.path.exists(LIBCPP):
        pytest.fail("libemu_cpp.so is not built (python -m emu_amd.build)")
    from emu_amd import capi
    capi.lib()  # loads libemu_amd.so first, from the package directory
    lib = ctypes.CDLL(LIBCPP)
    lib.emu_cpp_read_params.restype = ctypes.c_int
    lib.emu_cpp_read_params.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                        ctypes.POINTER(capi.EmuParams), ctypes.POINTER(capi.EmuInitConfig),
                                        ctypes.c_char_p, ctypes.c_int]
    return lib, capi


def read(cpp, path, overrides=()):
    lib, capi = cpp
    p, c = capi.EmuParams(), capi.EmuInitConfig()
    err = ctypes.create_string_buffer(512)
    arr = (ctypes.c_char_p * max(1, len(overrides)))(*[o.encode() for o in overrides])
    code = lib.emu_cpp_read_params(path.encode(), arr, len(overrides), ctypes.byref(p), ctypes.byref(c), err, 512)
    return code, p, c, err.value.decode()


def fields(s):
    out = {}
    for name, typ in s._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


@pytest.mark.parametrize("deck", sorted(json.load(open(os.path.join(ROOT, "tests", "golden", "reference_decks.json")))
                                        ["decks"]))
def test_cpp_reader_matches_python_reader(cpp, tmp_path, deck):
    """TestParams::Initialize (C++) and TestParams.from_file (Python) convert
    each of the reference's decks to identical parameter structs."""
    from emu_amd.params import TestParams
    src = os.path.join(ROOT, "tests", "golden", "reference_decks.json")
    d = json.load(open(src))["decks"][deck]
    path = tmp_path / "inputs"
    path.write_text(cases.deck_text(d))
    code, p, c, err = read(cpp, str(path))
    assert code == 0, err
    tp = TestParams.from_file(str(path))
    assert fields(p) == fields(tp.emu_params())
    assert fields(c) == fields(tp.init_config())


def test_cpp_reader_required_key_and_overrides(cpp, tmp_path):
    d = cases.fast_flavor_dict(nf=3)
    text = cases.deck_text(d)
    path = tmp_path / "inputs"
    path.write_text("\n".join(l for l in text.splitlines() if not l.startswith("nsteps")) + "\n")
    code, _, _, err = read(cpp, str(path))
    assert code != 0 and "nsteps" in err          # ParmParse::get aborts on a missing key
    path.write_text(text)
    code, p, c, err = read(cpp, str(path), ["ncell=(8,4,2)", "seed=7"])
    assert code == 0, err
    assert list(p.ncell) == [8, 4, 2] and c.seed == 7
    code, _, _, err = read(cpp, str(path), ["nf=4"])
    assert code != 0 and "nf" in err


def test_emu_main_print_params(tmp_path):
    d = cases.fast_flavor_dict(nf=2, ncell=(2, 3, 4))
    path = tmp_path / "inputs"
    path.write_text(cases.deck_text(d))
    out = subprocess.run([MAIN, str(path), "--print-params"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    p = json.loads(out.stdout)
    assert p["nf"] == 2 and p["ncell"] == [2, 3, 4]
    from emu_amd.params import TestParams
    tp = TestParams.from_file(str(path)).emu_params()
    assert p["mass"] == list(tp.mass) and p["theta12"] == tp.theta12


def test_cpp_library_exports(cpp):
    lib, _ = cpp
    for name in ("emu_cpp_read_params", "emu_cpp_evolve_flavor"):
        assert hasattr(lib, name)
