"""CPU restatement of the VelocityField / ShapeField box fields.

Test infrastructure only: the tests compare the library's rc_box_field with
these functions; nothing in the library imports them.

box_fields: VelocityField::calc_field (src/Observables/VelocityField.cpp:69-78)
and ShapeField::calc_field (src/Observables/ShapeField.cpp:69-78) add w * phi
to the field at the site's box index, over the fields in index order and
within a field over its sites q, with w the velocity row
(get_velocity_x/y) or the field's (S00, S01).  The box index of site q is
MultiPhaseField::GetSubIndex (src/Fields/MultiPhaseField.cpp:348-359): the
storage index q is the logical site ((q mod LX + offx) mod LX,
(q / LX + offy) mod LY) of the patch, placed at the corner, periodic.

coarse: the f_field_coarse loop of VelocityField::f_field
(VelocityField.cpp:106-121; ShapeField.cpp the same): site (j, i) adds
f(ss) / (g * g) over rows k then columns l, ss = ((j - g/2) mod L + l) mod L
+ L ((i - g/2) mod L + k) mod L).

Pinned by tests/test_box_field_oracle.py: g = 1 is the identity, a delta
spreads to the g x g block the index formula names with value exactly 1/g^2,
a constant field stays constant, the total of f is the sum of w * phi over all
rows, and a hand-placed 3 x 3 patch with offsets in a 5 x 5 box lands on the
sites worked out by hand.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np


def site_box_index(L: int, LX: int, LY: int, corner: int, offx: int, offy: int) -> np.ndarray:
    """Box index (x + L y) of every storage site q of a patch."""
    q = np.arange(LX * LY)
    qy, qx = q // LX, q % LX
    bx = ((qx + offx) % LX + corner % L) % L
    by = ((qy + offy) % LY + corner // L) % L
    return by * L + bx


def box_fields(L: int, geometry: Sequence[Tuple[int, int, int, int, int]], phis: Sequence[np.ndarray],
               weights: Sequence[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    """f_field_x, f_field_y (L x L, [y, x]).

    geometry: (LX, LY, corner, offx, offy) per field, in field order; phis: the
    fields in storage order; weights: per field either rows of shape (S, 2)
    (velocity) or a pair (S00, S01).
    """
    fx = np.zeros(L * L)
    fy = np.zeros(L * L)
    for (LX, LY, corner, offx, offy), phi, w in zip(geometry, phis, weights):
        k = site_box_index(L, LX, LY, corner, offx, offy)
        w = np.asarray(w, dtype=np.float64)
        if w.ndim == 2:
            wx, wy = w[:, 0], w[:, 1]
        else:
            wx, wy = np.full(phi.size, w[0]), np.full(phi.size, w[1])
        # a patch covers each box site at most once: one addition per site,
        # in field order, as the reference's loop makes them
        fx[k] += wx * phi
        fy[k] += wy * phi
    return fx.reshape(L, L), fy.reshape(L, L)


def coarse(f: np.ndarray, g: int) -> np.ndarray:
    """The g x g periodic average of f (L x L, [y, x]) in the reference's order."""
    out = np.zeros_like(f)
    h = g // 2
    for k in range(g):
        for l in range(g):
            out += np.roll(f, shift=(-(k - h), -(l - h)), axis=(0, 1)) / float(g * g)
    return out
