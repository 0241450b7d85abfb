"""numpy restatement of the force part of the 2D CALCCOEF (cylinder/cylinder_v3.0.3_AF_original
4step_face_nl/calccoef.f90:46-119): per wall cell J, the wall distance DN1 and the wall angle
TH2 from the node triangle, du_t/dn = (u sin TH2 - v cos TH2) / DN1, and the pressure and
friction drag and lift summed in J order.  Test infrastructure (pinned against the flang
reference's scalars in tests/test_geometry2d.py)."""
import math

import numpy as np


def calccoef2d(xg, yg, x, y, theta, u, v, p, re):
    """xg, yg: (NI, NJ) nodes; x, y, theta, u, v, p: (NI+1, NJ+1) cell arrays."""
    nj = xg.shape[1]
    pi = 4.0 * math.atan(1.0)
    pd = pl = td = tl = 0.0
    for j in range(1, nj):
        X, Y = x[1, j], y[1, j]
        xa, ya, xb, yb = xg[0, j - 1], yg[0, j - 1], xg[0, j], yg[0, j]
        l1 = ((X - xa) ** 2.0 + (Y - ya) ** 2.0) ** 0.5
        l2 = ((X - xb) ** 2.0 + (Y - yb) ** 2.0) ** 0.5
        l3 = ((xb - xa) ** 2.0 + (yb - ya) ** 2.0) ** 0.5
        th1 = math.acos((l1 * l1 + l2 * l2 - l3 * l3) / (2.0 * l1 * l2))
        dn1 = l1 * l2 / l3 * math.sin(th1)
        la = (yb - ya) / (xb - xa)
        lb = ya - la * xa
        la_ = -1.0 / la
        lb_ = Y - la_ * X
        x1 = (lb_ - lb) / (la - la_)
        y1 = la * (lb_ - lb) / (la - la_) + lb
        dn2 = ((Y - y1) ** 2.0 + (X - x1) ** 2.0) ** 0.5
        th2 = math.atan((Y - y1) / (X - x1))
        a = math.asin((Y - y1) / dn2)
        if th2 >= 0.0 and a > 0.0:
            pass
        elif th2 < 0.0 and a > 0.0:
            th2 = pi + th2
        elif th2 >= 0.0 and a <= 0.0:
            th2 = pi + th2
        elif th2 < 0.0 and a < 0.0:
            th2 = 2.0 * pi + th2
        dudn = (u[1, j] * math.sin(th2) - v[1, j] * math.cos(th2) - 0.0) / dn1
        ct, st = math.cos(theta[1, j]), math.sin(theta[1, j])
        pd = pd - 2.0 * p[1, j] * l3 * ct
        pl = pl - 2.0 * p[1, j] * l3 * st
        td = td + (2.0 / re) * dudn * l3 * st
        tl = tl - (2.0 / re) * dudn * l3 * ct
    return dict(cd=pd + td, cl=pl + tl, cd_p=pd, cd_v=td, cl_p=pl, cl_v=tl,
                cpb=2.0 * (0.0 - 0.25 * (p[1, 1] + p[1, 0] + p[0, 1] + p[0, 0])))
