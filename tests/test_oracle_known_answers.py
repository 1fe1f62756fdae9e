"""Known answers from the reference's own unit tests, held against the oracle (CPU only).

* src/test/test_pbc.cpp:13-26: the BvK cell list of a 2x2x2 period is ordered x, y, z from
  -(p / 2) (construct_R_grid, src/core/pbc.cpp:178-190), so R = (-1, -1, -1) is cell 0 and
  (-1, 0, -1) is cell 2; (3, 3, -1) is no cell of the list but folds to cell 0 under the
  symmetric modulo (vector3_order.h:75-80).
* src/test/test_fitting.cpp:9-40: the Levenberg-Marquardt fit of sin(a x) on 100 points of
  [-pi/2, pi/2], started at a = 1, recovers a = 0.5, 1.5, 2 and -2 to 1e-8. The oracle's
  ``levmarq_fit`` restates the reference's fit (src/math/fitting.cpp:80-182) that the
  dielectric-function option 2 uses.
"""
import math

import numpy as np

from oracle import dielec, rpa


def test_bvk_cell_order_and_fold():
    cells = rpa.r_grid((2, 2, 2))
    assert len(cells) == 8
    assert cells.index((-1, -1, -1)) == 0
    assert cells.index((-1, 0, -1)) == 2
    assert (3, 3, -1) not in cells
    assert cells.index(rpa.fold_R((3, 3, -1), (2, 2, 2))) == 0


def test_levmarq_recovers_sine_frequency():
    n = 100
    xs = [(i / (n - 1) - 0.5) * math.pi for i in range(n)]
    f = lambda x, p: math.sin(x * p[0])
    g = lambda x, p: [x * math.cos(x * p[0])]
    for a in (0.5, 1.5, 2.0, -2.0):
        ys = [math.sin(a * x) for x in xs]
        p = dielec.levmarq_fit([1.0], xs, ys, f, g)
        assert abs(p[0] - a) <= 1e-8, (a, p)
    # a plausible mistake, the damping moved the wrong way after a rejected step, must not pass
    # unnoticed: the same fit with the up and down factors swapped ends elsewhere for some a
    wrong = [dielec.levmarq_fit([1.0], xs, [math.sin(a * x) for x in xs], f, g, up=0.1, down=0.1)[0]
             for a in (0.5, 1.5, 2.0, -2.0)]
    assert not np.allclose(wrong, [0.5, 1.5, 2.0, -2.0], atol=1e-8)
