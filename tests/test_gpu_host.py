"""The C++ drop-in executable (FD_HIPBackend + SimManager) end to end on the
GPU, against the oracle run of the same input."""
import os
import subprocess

import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd.conf import read_conf
from tests.host_case import make_case

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "rheocell_amd", "rheoCell_hip")


@pytest.mark.parametrize("fc", [0, 1], ids=["dry", "wet"])
def test_executable_matches_oracle(tmp_path, fc):
    conf, keys = make_case(str(tmp_path), L=64, nside=3, seed=5, steps=20, print_every=10, friction_cell=fc,
                           CGtolerance=1e-13)
    r = subprocess.run([EXE, "input"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    last = read_conf(str(tmp_path / "last_conf.dat"), len(conf.cells))
    assert last.step == 20
    orc = OracleSim(conf, keys)
    orc.step(20)
    for i, c in enumerate(last.cells):
        oc = orc.cell(i)
        assert (c.LsubX, c.LsubY, c.offset, c.corner) == (oc.LsubX, oc.LsubY, tuple(oc.offset), oc.corner)
        assert np.abs(c.phi - orc.phi(i)).max() < 1e-10
        assert np.array_equal(c.box_index, orc.site(i, 8).astype(np.int64))
        assert abs(c.CoM[0] - oc.CoM[0]) < 1e-9 and abs(c.Q00 - oc.Q00) < 1e-12
    traj = open(tmp_path / "trajectory.dat").read()
    assert traj.count("t = ") == 2  # steps 10 and 20


@pytest.mark.parametrize("interaction", ["wetpolarmodel", "wetmodel"])
def test_executable_with_walls_and_external_forces(tmp_path, interaction):
    # box = square with type = square_walls, and an external_forces_file of the
    # reference's format (brace blocks, # comments)
    from rheocell_amd.walls import read_forces, wall_field
    forces = ("# walls and a shear flow\n{\ntype = square_walls\nfield = -1\nkappa_wall = 0.2\n}\n"
              "{\ntype = shear_flow_channel\nfield = -1\nkappa_wall = 0.1\nlambda = 3.0\nshear_rate = 0.05\n}\n")
    (tmp_path / "external.conf").write_text(forces)
    extra = dict(interaction_type="wetmodel", J_Q=0.1, zetaQ_self=0.3) if interaction == "wetmodel" else {}
    conf, keys = make_case(str(tmp_path), L=96, nside=4, seed=6, steps=12, print_every=12, type="square_walls",
                           lambda_wall=4, external_forces=1, external_forces_file="external.conf", **extra)
    r = subprocess.run([EXE, "input"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    last = read_conf(str(tmp_path / "last_conf.dat"), len(conf.cells))
    orc = OracleSim(conf, keys, walls=wall_field("square_walls", 96, 4.0), forces=read_forces(forces))
    orc.step(12)
    for i, c in enumerate(last.cells):
        oc = orc.cell(i)
        assert (c.LsubX, c.LsubY, c.offset, c.corner) == (oc.LsubX, oc.LsubY, tuple(oc.offset), oc.corner)
        assert np.abs(c.phi - orc.phi(i)).max() < 1e-10


def test_executable_runs_the_reference_example_input(tmp_path):
    # examples/input of the reference (wetmodel, 100 cells at 140^2), with a
    # conf of 100 cells, no equilibration, 4 steps and a tight solver
    # tolerance so that the oracle's solve is the same to 1e-10
    import json
    ex = json.load(open(os.path.join(ROOT, "tests", "golden", "example_input.json")))["keys"]
    ex = dict(ex, backend="HIP", equilibration_steps=0, steps=4, print_conf_interval=4, CGtolerance=1e-12)
    from rheocell_amd import confgen
    from rheocell_amd.conf import write_conf, write_input, write_topology
    conf = confgen.lattice_conf(140, 10, seed=1)
    with open(tmp_path / "start_conf.dat", "w") as f:
        write_conf(conf, f)
    write_topology(str(tmp_path / "topology.top"), 100)
    write_input(str(tmp_path / "input"), ex)
    r = subprocess.run([EXE, "input"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    last = read_conf(str(tmp_path / "last_conf.dat"), 100)
    orc = OracleSim(conf, ex)
    orc.step(4)
    for i, c in enumerate(last.cells):
        assert np.abs(c.phi - orc.phi(i)).max() < 1e-10


@pytest.mark.parametrize("restart", [1, 0])
def test_trajectory_restart_and_output_prefix(tmp_path, restart):
    # restart_step_counter = 1 starts the trajectory afresh, 0 appends to it
    # (ObservableOutput.cpp:31-33, 94-99); output_prefix prefixes every name
    make_case(str(tmp_path), L=64, nside=2, seed=3, steps=4, print_every=2, output_prefix="out_",
              restart_step_counter=restart)
    stale = "t = 999\nstale frame\n"
    (tmp_path / "out_trajectory.dat").write_text(stale)
    r = subprocess.run([EXE, "input"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    traj = (tmp_path / "out_trajectory.dat").read_text()
    assert traj.startswith(stale) == (restart == 0)
    assert traj.count("t = ") == 2 + (restart == 0)
    assert (tmp_path / "out_last_conf.dat").exists() and not (tmp_path / "trajectory.dat").exists()


def test_async_conf_output_matches_inline(tmp_path):
    # the configuration writer thread produces the same files, byte for byte,
    # as writing inline (async_conf_output = 0), at every print of the run
    outs = {}
    for mode in (1, 0):
        d = tmp_path / ("async%d" % mode)
        make_case(str(d), L=64, nside=3, seed=9, steps=12, print_every=2, async_conf_output=mode)
        r = subprocess.run([EXE, "input"], cwd=str(d), capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs[mode] = (open(d / "trajectory.dat", "rb").read(), open(d / "last_conf.dat", "rb").read())
    assert outs[1] == outs[0]
    assert outs[1][0].count(b"t = ") == 6
