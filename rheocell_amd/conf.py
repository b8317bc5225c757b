"""Configuration (conf / trajectory frame), topology and input-file formats.

Byte-level format of the reference:

* conf frame -- ``Configuration::headers`` / ``Configuration::field``
  (src/Observables/Configuration.cpp:36-55): ``t = <step>`` line, ``b = Lx Ly``
  line, then one line per cell ``LsubX LsubY CoMx CoMy offx offy corner
  unrap_x unrap_y Q00 Q01`` followed by ``box_index phi`` pairs in storage
  order, numbers printed with 15 significant digits and a trailing blank.
* reader -- ``SimBackend::read_next_configuration``
  (src/Backends/SimBackend.cpp:220-314): header errors and a cell-count
  mismatch raise.
* topology -- first token is the number of cells
  (src/Interactions/BaseInteraction.cpp:27-35).
* input file -- ``key = value`` lines, ``#`` comments, command-line
  ``key=value`` overrides (src/Utilities/parse_input/parse_input.cpp:20-156).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, Iterable, List, Optional, TextIO

import numpy as np


class RCexception(RuntimeError):
    """Mirror of the reference's ``RCexception`` (src/Utilities/RCexception.h)."""


@dataclasses.dataclass
class CellConf:
    LsubX: int
    LsubY: int
    CoM: tuple
    offset: tuple
    corner: int
    unrap: tuple
    Q00: float
    Q01: float
    box_index: np.ndarray  # int64, storage order
    phi: np.ndarray        # float64, storage order

    @property
    def subSize(self) -> int:
        return self.LsubX * self.LsubY


@dataclasses.dataclass
class Conf:
    step: int
    Lx: int
    Ly: int
    cells: List[CellConf]


def fmt15(x) -> str:
    """C++ ostream with precision(15), default float field: ``%.15g``."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    return "%.15g" % float(x)


def format_cell(c: CellConf) -> str:
    head = [c.LsubX, c.LsubY, float(c.CoM[0]), float(c.CoM[1]), int(c.offset[0]), int(c.offset[1]),
            int(c.corner), float(c.unrap[0]), float(c.unrap[1]), float(c.Q00), float(c.Q01)]
    parts = [fmt15(v) for v in head]
    body = []
    for k, v in zip(c.box_index.tolist(), c.phi.tolist()):
        body.append(str(int(k)))
        body.append("%.15g" % v)
    return " ".join(parts) + " " + " ".join(body) + " "


def write_conf(conf: Conf, f: TextIO) -> None:
    f.write("t = %d\n" % conf.step)
    f.write("b = %d %d\n" % (conf.Lx, conf.Ly))
    f.write("\n".join(format_cell(c) for c in conf.cells))


def _read_frame(lines: Iterable[str], ncells: int) -> Optional[Conf]:
    it = iter(lines)
    try:
        l0 = next(it)
    except StopIteration:
        return None
    l0 = l0.rstrip("\n")
    if not l0.startswith("t = "):
        raise RCexception('Malformed headers found in an input configuration."t = <int>" was expected, but "%s" was found instead.' % l0)
    step = int(l0[4:].split()[0])
    try:
        l1 = next(it).rstrip("\n")
    except StopIteration:
        raise RCexception("Malformed headers: missing box line")
    if not l1.startswith("b = "):
        raise RCexception('Malformed headers found in an input configuration."b = <float> <float> <float>" was expected, but "%s" was found instead.' % l1)
    bx = l1[4:].split()
    Lx, Ly = int(float(bx[0])), int(float(bx[1]))
    cells = []
    for _ in range(ncells):
        try:
            line = next(it)
        except StopIteration:
            break
        vals = line.split()
        if not vals:
            break
        LX, LY = int(float(vals[0])), int(float(vals[1]))
        S = LX * LY
        body = np.array(vals[11:11 + 2 * S], dtype=np.float64)
        if body.size != 2 * S:
            raise RCexception("configuration line for cell %d is truncated" % len(cells))
        cells.append(CellConf(LX, LY, (float(vals[2]), float(vals[3])), (int(float(vals[4])), int(float(vals[5]))),
                              int(float(vals[6])), (float(vals[7]), float(vals[8])), float(vals[9]), float(vals[10]),
                              body[0::2].astype(np.int64), body[1::2].copy()))
    if len(cells) != ncells:
        raise RCexception("The number of lines found in configuration file (%d) doesn't match the parsed number of particles (%d)" % (len(cells), ncells))
    return Conf(step, Lx, Ly, cells)


def read_conf(path: str, ncells: int) -> Conf:
    with open(path) as f:
        conf = _read_frame(f, ncells)
    if conf is None:
        raise RCexception("Could not read the initial configuration, aborting")
    return conf


def read_topology(path: str) -> int:
    try:
        with open(path) as f:
            tok = f.read().split()
    except OSError:
        raise RCexception("Can't read topology file '%s'. Aborting" % path)
    return int(tok[0])


def write_topology(path: str, n: int) -> None:
    with open(path, "w") as f:
        f.write("%d\n" % n)


def parse_input(text: str, overrides: Iterable[str] = ()) -> Dict[str, str]:
    """``key = value`` subset of the reference input format."""
    out: Dict[str, str] = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        k, v = line.split("=", 1)
        out[k.strip()] = v.strip()
    for ov in overrides:
        k, v = ov.split("=", 1)
        out[k.strip()] = v.strip()
    return out


def read_input(path: str, overrides: Iterable[str] = ()) -> Dict[str, str]:
    with open(path) as f:
        return parse_input(f.read(), overrides)


def write_input(path: str, keys: Dict[str, object]) -> None:
    with open(path, "w") as f:
        for k, v in keys.items():
            f.write("%s = %s\n" % (k, v))
