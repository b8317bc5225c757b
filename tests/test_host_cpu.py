"""CPU tests of the boundary: the C ABI exports, the C++ drop-in's input /
configuration handling and its error behaviour (all before any GPU call)."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.host_case import make_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "rheocell_amd", "rheoCell_hip")
LIB = os.path.join(ROOT, "rheocell_amd", "librheocell.so")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "rheocell.h")).read()
    names = re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(rc_\w+)\s*\(", hdr, re.M)
    assert len(names) >= 15
    lib = ctypes.CDLL(LIB)
    for n in names:
        assert hasattr(lib, n), n
    from rheocell_amd import _lib
    assert set(_lib.EXPORTS) <= set(names)


def _run(d, *over):
    return subprocess.run([EXE, "input"] + list(over), cwd=d, capture_output=True, text=True, timeout=60)


def test_unknown_backend_is_rejected(tmp_path):
    make_case(str(tmp_path))
    r = _run(str(tmp_path), "backend=CPU")
    assert r.returncode == 1 and "backend `CPU' is not provided" in r.stderr


def test_missing_mandatory_key(tmp_path):
    make_case(str(tmp_path))
    txt = open(tmp_path / "input").read().replace("dt = 0.1\n", "")
    open(tmp_path / "input", "w").write(txt)
    r = _run(str(tmp_path))
    assert r.returncode == 1 and "Mandatory key `dt' not found" in r.stderr


def test_uninitialised_reference_keys_are_mandatory(tmp_path):
    make_case(str(tmp_path))
    txt = open(tmp_path / "input").read().replace("J_Q = 0\n", "")
    open(tmp_path / "input", "w").write(txt)
    r = _run(str(tmp_path))
    assert r.returncode == 1 and "Mandatory key `J_Q' not found" in r.stderr


def test_malformed_conf_header(tmp_path):
    make_case(str(tmp_path))
    lines = open(tmp_path / "start_conf.dat").read().split("\n")
    lines[0] = "x = 0"
    open(tmp_path / "start_conf.dat", "w").write("\n".join(lines))
    r = _run(str(tmp_path))
    assert r.returncode == 1 and "Malformed headers" in r.stderr


def test_non_square_box(tmp_path):
    make_case(str(tmp_path))
    lines = open(tmp_path / "start_conf.dat").read().split("\n")
    lines[1] = "b = 64 80"
    open(tmp_path / "start_conf.dat", "w").write("\n".join(lines))
    r = _run(str(tmp_path))
    assert r.returncode == 1 and "not a square" in r.stderr


def test_cell_count_mismatch(tmp_path):
    make_case(str(tmp_path))
    open(tmp_path / "topology.top", "w").write("5\n")
    r = _run(str(tmp_path))
    assert r.returncode == 1 and "doesn't match the parsed number of particles" in r.stderr


def test_equilibration_requires_restart_counter(tmp_path):
    make_case(str(tmp_path))
    r = _run(str(tmp_path), "equilibration_steps=10", "restart_step_counter=0")
    assert r.returncode == 1 and "restart_step_counter must be set to true" in r.stderr


def test_conf_round_trip(tmp_path):
    from rheocell_amd.conf import read_conf, write_conf
    conf, _ = make_case(str(tmp_path))
    back = read_conf(str(tmp_path / "start_conf.dat"), len(conf.cells))
    with open(tmp_path / "again.dat", "w") as f:
        write_conf(back, f)
    assert open(tmp_path / "again.dat").read() == open(tmp_path / "start_conf.dat").read()


REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="the reference sources are not present")
def test_fd_hip_backend_compiles_against_the_reference_headers(tmp_path):
    # the in-tree binding (FD_HIPBackend : FDBackend) against the reference's
    # own SimBackend / FDBackend / BaseField / BaseInteraction headers; the
    # instance proves every pure virtual is implemented
    src = tmp_path / "check.cpp"
    src.write_text('#include "FD_HIPBackend.h"\nvoid make_backend() { FD_HIPBackend b; (void) b; }\n')
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-iquote", os.path.join(REF_SRC, "Backends"),
                        "-I" + REF_SRC, "-I" + os.path.join(REF_SRC, "extern"), "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "rheocell_amd", "host"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_reference_example_input_is_accepted():
    # the key = value pairs of the reference's examples/input (a fixture of
    # tests/golden/make_golden.py) map to the nematic model's parameters
    import json
    from rheocell_amd import _lib
    from rheocell_amd.sim import model_params
    keys = json.load(open(os.path.join(ROOT, "tests", "golden", "example_input.json")))["keys"]
    p = model_params(keys)
    assert p.interaction == _lib.RC_INTERACTION["wetmodel"]
    assert (p.gamma, p.mu, p.lambda_, p.kappa, p.wall_slip, p.friction_cell, p.CGtolerance) == \
        (0.06, 20.0, 2.0, 0.5, 2.0, 1.0, 1e-5)
    assert (p.J_Q, p.zetaQ_self, p.zetaQ_inter, p.beta) == (0.1, 0.3, 0.0, 0.0)
