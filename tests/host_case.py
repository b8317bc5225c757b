"""Writes a runnable input directory (input, topology, start conf) for the
C++ drop-in executable."""
import os

from rheocell_amd import confgen
from rheocell_amd.conf import write_conf, write_input, write_topology


def make_case(d, L=64, nside=2, seed=7, steps=20, print_every=10, **over):
    conf = confgen.lattice_conf(L, nside, seed)
    keys = dict(confgen.BASE_KEYS, seed=seed, steps=steps, print_conf_interval=print_every)
    keys.update(over)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "start_conf.dat"), "w") as f:
        write_conf(conf, f)
    write_topology(os.path.join(d, "topology.top"), len(conf.cells))
    write_input(os.path.join(d, "input"), keys)
    return conf, keys
