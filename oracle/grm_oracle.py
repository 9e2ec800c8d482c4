"""CPU restatement of `pair_style lubricate/GRM/LBDEM` (test infrastructure).

Only the tests import this module: it is the checker of the C++ pair style
in lammps-lbm-psm_amd/fix/pair_GRM_LBDEM.cpp, never a product path.  It
restates PairLubricateGRMLBDEM::compute of src/pair_GRM_LBDEM.cpp:78-476
in plain Python floats (IEEE doubles, math.log and math.sqrt of the C
library), one pair at a time, in the reference's operation order:

* gap = |x_j - x_i| - a_i - a_j; outer gap g_N = sqrt(cut_ij^2) * cutoff_factor
  (:160-167); pairs with gap >= g_N are skipped (:169);
* velocities enter negated (:140-148, :173-179); a gap below cut_inner
  takes cut_inner (:182-183); both gaps are scaled by the mean radius
  (:186-197);
* one resistance term per flagGRM (:258-306), evaluated at the gap and at
  g_N, and the component of the relative motion it acts on (:325-384);
* force (flags 1, 2, 4) or torque (flags 3, 5) of atom i only, times
  force->vxmu2f (:388-453), and the virial tallied for i alone with
  del = x_i - x_j (:457-458).

How this restatement is pinned (tests/test_pair_grm.py): the squeeze force
of two equal spheres against Reynolds' lubrication force 3 pi mu a^2 V / (2 h)
and of a small sphere near a much larger one against the sphere-plane force
6 pi mu a^2 V / h; the shear term of equal spheres against Jeffrey & Onishi's
(1/6) ln(1/xi) coefficient of 6 pi mu a; reciprocity of the rotation-
translation coupling (terms 3 and 4, B_11 = Bt_11); equal and opposite pair
forces for equal radii (flags 1 and 2); the sign of the pair virial; no
force for a common translation; a force that vanishes continuously at the
outer gap and saturates below the inner one.  Each plausible slip tried on
it (an exponent, a sign of the velocities or the virial, ln(h) for ln(1/h),
a swapped cross product, a coefficient) fails one of these checks.
"""
from __future__ import annotations

import math

MY_PI = 3.14159265358979323846


def resistance(flag, beta, ai, aj, h):
    """(self, partner) coefficients of term `flag` at scaled gap h,
    src/pair_GRM_LBDEM.cpp:228-306."""
    b0p2 = beta * beta
    b1 = 1.0 + beta
    b1p2 = b1 * b1
    b1p3 = b1p2 * b1
    lg = math.log(1.0 / h)
    if flag == 1:
        x11 = 2.0 * b0p2 / b1p3 / h
        x11 *= 6.0 * MY_PI * ai
        return x11, -x11
    if flag == 2:
        y11 = 4.0 * beta * (2.0 + beta + 2.0 * b0p2) / (15.0 * b1p3) * lg
        y11 *= 6.0 * MY_PI * ai
        return y11, -y11
    if flag == 3:
        y11 = -beta * (4.0 + beta) / (5.0 * b1p2) * lg
        y11 *= 4.0 * MY_PI * (ai * ai)
        return y11, -y11
    if flag == 4:
        y11 = -beta * (4.0 + beta) / (5.0 * b1p2) * lg
        y11 *= 4.0 * MY_PI * (ai * ai)
        bi = 1.0 / beta
        bi1 = 1.0 + bi
        y21 = bi * (4.0 + bi) / (5.0 * (bi1 * bi1)) * lg
        y21 *= -4.0 * MY_PI * (aj * aj)
        return y11, y21
    if flag == 5:
        ai3 = (ai * ai) * ai
        y11 = 2.0 * beta / (5.0 * b1) * lg
        y11 *= 8.0 * MY_PI * ai3
        y12 = b0p2 / (10.0 * b1) * lg
        y12 *= 8.0 * MY_PI * ai3
        return y11, y12
    return 0.0, 0.0


def _dot(a, e):
    return a[0] * e[0] + a[1] * e[1] + a[2] * e[2]


def compute(x, v, omega, radius, types, nlocal, neigh, *, mu, flag, cut, cut_inner,
            cutoff_factor, vxmu2f=1.0):
    """Forces, torques (nlocal x 3 lists) and global virial (6) of the pair.

    x, v, omega: sequences of 3-vectors; radius, types: per atom; neigh[i]:
    the full-list partners of local i in order; cut, cut_inner: dicts keyed
    by (itype, jtype), symmetric.
    """
    f = [[0.0, 0.0, 0.0] for _ in range(nlocal)]
    t = [[0.0, 0.0, 0.0] for _ in range(nlocal)]
    vir = [0.0] * 6
    for i in range(nlocal):
        ai = radius[i]
        vi = [-v[i][0], -v[i][1], -v[i][2]]
        wi = [-omega[i][0], -omega[i][1], -omega[i][2]]
        for j in neigh[i]:
            ti, tj = types[i], types[j]
            delx = x[j][0] - x[i][0]
            dely = x[j][1] - x[i][1]
            delz = x[j][2] - x[i][2]
            rsq = delx * delx + dely * dely + delz * delz
            aj = radius[j]
            c = cut[(ti, tj)]
            gap_n = math.sqrt(c * c) * cutoff_factor
            r = math.sqrt(rsq)
            gap = r - ai - aj
            if not gap < gap_n:
                continue
            vj = [-v[j][0], -v[j][1], -v[j][2]]
            wj = [-omega[j][0], -omega[j][1], -omega[j][2]]
            if gap < cut_inner[(ti, tj)]:
                gap = cut_inner[(ti, tj)]
            am = 0.5 * (ai + aj)
            h = gap / am
            hn = gap_n / am
            beta = aj / ai
            s11, s12 = resistance(flag, beta, ai, aj, h)
            n11, n12 = resistance(flag, beta, ai, aj, hn)
            e = [delx / r, dely / r, delz / r]
            if flag in (1, 2):
                vin, vjn = _dot(vi, e), _dot(vj, e)
                si = [vin * e[k] for k in range(3)]
                sj = [vjn * e[k] for k in range(3)]
                if flag == 2:
                    si = [vi[k] - si[k] for k in range(3)]
                    sj = [vj[k] - sj[k] for k in range(3)]
            elif flag == 3:
                si = [vi[1] * e[2] - vi[2] * e[1], vi[2] * e[0] - vi[0] * e[2],
                      vi[0] * e[1] - vi[1] * e[0]]
                sj = [vj[1] * e[2] - vj[2] * e[1], vj[2] * e[0] - vj[0] * e[2],
                      vj[0] * e[1] - vj[1] * e[0]]
            elif flag == 4:
                si = [e[1] * wi[2] - e[2] * wi[1], e[2] * wi[0] - e[0] * wi[2],
                      e[0] * wi[1] - e[1] * wi[0]]
                sj = [e[1] * wj[2] - e[2] * wj[1], e[2] * wj[0] - e[0] * wj[2],
                      e[0] * wj[1] - e[1] * wj[0]]
            elif flag == 5:
                win, wjn = _dot(wi, e), _dot(wj, e)
                si = [wi[k] - win * e[k] for k in range(3)]
                sj = [wj[k] - wjn * e[k] for k in range(3)]
            else:
                si = [0.0, 0.0, 0.0]
                sj = [0.0, 0.0, 0.0]
            out = [mu * ((s11 * si[k] + s12 * sj[k]) - (n11 * si[k] + n12 * sj[k]))
                   for k in range(3)]
            fx = fy = fz = 0.0
            if flag in (1, 2, 4):
                fx, fy, fz = out[0] * vxmu2f, out[1] * vxmu2f, out[2] * vxmu2f
                f[i][0] += fx
                f[i][1] += fy
                f[i][2] += fz
            elif flag in (3, 5):
                for k in range(3):
                    t[i][k] += vxmu2f * out[k]
            dx, dy, dz = -delx, -dely, -delz
            pv = [dx * fx, dy * fy, dz * fz, dx * fy, dx * fz, dy * fz]
            for k in range(6):
                vir[k] += 0.5 * pv[k]
    return f, t, vir


def mix_geometric(a, b):
    """Pair::mix_distance for the default mix_flag (geometric)."""
    return math.sqrt(a * b)
