"""Zou-He pins of the CPU oracle (no GPU needed).

Every face function of src/LBM_PSM_zou_he_BC.cpp sets the populations that
enter the domain through the face (those with e . n > 0, n the inward
normal) and leaves the others.  The Zou-He construction (Zou & He 1997;
Hecht & Harting 2010 for D3Q19) fixes the unknowns by four conditions, which
are checked here on random non-equilibrium states with random imposed
velocities, independently of how the formulas are written:

1. the known populations are unchanged;
2. velocity face: the node's momentum after the face is rho u_imposed, with
   rho the node's own density after the face (this also fixes rho to the
   known-population expression (K_0 + 2 K_out) / (1 - u.n));
   density face: the density is the imposed one and the tangential momentum
   is rho u_t (u_t = 0 in 2D);
3. the non-equilibrium part of the population along the normal bounces
   back: f_n - f_-n = feq_n - feq_-n = 6 w_n rho (e_n . u);
4. (D3Q19) the same bounce-back holds for the sum over each pair of
   unknown diagonals that share a tangential axis, i.e. the transverse
   corrections only move momentum within a pair.

In D2Q9 conditions 2-3 give three equations for the three unknowns and in
D3Q19 conditions 2-4 give five for five, so they determine every unknown:
a wrong coefficient, sign or index in any face formula fails one of them.

The keyword tests then check the face sets per `lbm-psm-bc` keyword
(src/fix_LBM_PSM_BC.cpp:274-400): on every face, the nodes that no later
face of the keyword touches satisfy that face's conditions with the
keyword's velocity (signs included), and nodes shared with a later face
satisfy the later one (later faces overwrite edges and corners).
"""
import numpy as np
import pytest

import lbmpsm_oracle as O
from test_oracle_pins import D2Q9, D3Q19

TOL = 2e-14

# kind -> (dim, normal axis, inward normal sign, density face)
FACES = {
    O_k: v for O_k, v in [
        (0, (2, 0, +1, False)),   # zh2_vel_xn,  :25-40
        (1, (2, 0, -1, False)),   # zh2_vel_xp,  :43-58
        (2, (2, 0, -1, True)),    # zh2_dens_xp, :61-76
        (3, (2, 1, +1, False)),   # zh2_vel_yn,  :79-94
        (4, (2, 1, -1, False)),   # zh2_vel_yp,  :97-112
        (10, (3, 0, +1, False)),  # zh3_vel_xn,  :153-185
        (11, (3, 0, -1, False)),  # zh3_vel_xp,  :189-221
        (12, (3, 0, -1, True)),   # zh3_dens_xp, :225-257
        (13, (3, 1, +1, False)),  # zh3_vel_yn,  :261-293
        (14, (3, 1, -1, False)),  # zh3_vel_yp,  :296-328
        (15, (3, 2, +1, False)),  # zh3_vel_zn,  :331-363
        (16, (3, 2, -1, False)),  # zh3_vel_zp,  :366-398
    ]
}
NAMES = {0: "2d-vel-xn", 1: "2d-vel-xp", 2: "2d-dens-xp", 3: "2d-vel-yn", 4: "2d-vel-yp",
         10: "3d-vel-xn", 11: "3d-vel-xp", 12: "3d-dens-xp", 13: "3d-vel-yn", 14: "3d-vel-yp",
         15: "3d-vel-zn", 16: "3d-vel-zp"}


def random_state(o, lat, rng):
    """Positive populations far from equilibrium: a random equilibrium per
    node plus a 5 % random non-equilibrium part."""
    e = lat["e"].astype(float)
    w = lat["w"]
    nh = o.nh
    rho = rng.uniform(0.9, 1.1, nh)
    u = rng.uniform(-0.04, 0.04, nh + (3,))
    if o.dim == 2:
        u[..., 2] = 0.0
    eu = u @ e.T
    usq = np.sum(u * u, axis=-1)[..., None]
    feq = rho[..., None] * w * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * usq)
    f = feq * (1 + 0.05 * rng.standard_normal(feq.shape))
    o.set_f_halo(f)
    return f


def check_face_nodes(lat, before, after, axis, sign, u_imp=None, rho_imp=None, ut=(0.0, 0.0, 0.0),
                     dim=3, rho_star=None):
    """before/after: [nodes, q] populations of the nodes of one face.
    rho_star: the density the face builds its unknowns with, where it is not
    the node's density after the face (see REFERENCE_DENSITY)."""
    e = lat["e"]
    w = lat["w"]
    q = len(w)
    nrm = np.zeros(3, int)
    nrm[axis] = sign
    s = e @ nrm
    unknown = s > 0
    # 1. the known populations are unchanged
    if before is not None:
        assert np.array_equal(after[:, ~unknown], before[:, ~unknown])
    rho = after.sum(axis=1) if rho_star is None else rho_star
    j = after @ e.astype(float)
    tang = [a for a in range(dim) if a != axis]
    if rho_imp is None:
        # 2. velocity face: j = rho u_imposed
        u = np.broadcast_to(np.asarray(u_imp, float), j.shape)
        assert np.max(np.abs(j - rho[:, None] * u)) < TOL
    else:
        # 2. density face: rho and the tangential momentum
        assert np.max(np.abs(rho - rho_imp)) < TOL
        for a in tang:
            assert np.max(np.abs(j[:, a] - rho * ut[a])) < TOL
        u = j / rho[:, None]
    opp = [int(np.flatnonzero(np.all(e == -e[i], axis=1))[0]) for i in range(q)]
    # 3. normal non-equilibrium bounce-back
    n_i = int(np.flatnonzero(np.all(e == nrm, axis=1))[0])
    lhs = after[:, n_i] - after[:, opp[n_i]]
    rhs = 6.0 * w[n_i] * rho * (u @ e[n_i])
    assert np.max(np.abs(lhs - rhs)) < TOL
    # 4. per-pair bounce-back of the unknown diagonals (D3Q19)
    if dim == 3:
        for a in tang:
            pair = [i for i in range(q) if unknown[i] and e[i, a] != 0]
            assert len(pair) == 2
            lhs = sum(after[:, i] - after[:, opp[i]] for i in pair)
            rhs = sum(6.0 * w[i] * rho * (u @ e[i]) for i in pair)
            assert np.max(np.abs(lhs - rhs)) < TOL


# The reference's 2D velocity xp face (src/LBM_PSM_zou_he_BC.cpp:47-48)
# takes rho = (K_0 + 2 K_out) / (1 - u) where Zou-He's density at an
# outflow face is (K_0 + 2 K_out) / (1 + u); it builds the unknowns from that
# density exactly as Zou-He does.  `lbm-psm-bc` reaches this face only in
# closedBox, with u = 0, where the two agree.  For u != 0 the oracle keeps
# the reference's form, and the test checks conditions 2-3 with that density
# (the node then carries momentum rho_star u and density rho_star (1 - 2u)).
REFERENCE_DENSITY = {1: lambda b, u: (b[:, 0] + b[:, 2] + b[:, 4]
                                      + 2.0 * (b[:, 1] + b[:, 5] + b[:, 8])) / (1.0 - u[0])}


def face_nodes(f, axis, plane):
    sl = [slice(None)] * 3
    sl[axis] = plane
    return f[tuple(sl)].reshape(-1, f.shape[-1])


@pytest.mark.parametrize("kind", sorted(FACES), ids=[NAMES[k] for k in sorted(FACES)])
def test_face_formula_meets_zou_he_conditions(kind):
    dim, axis, sign, dens = FACES[kind]
    lat = D3Q19 if dim == 3 else D2Q9
    rng = np.random.default_rng(100 + kind)
    o = O.Oracle(dim, (5, 6, 7), (1, 1, 1), 0.7)
    nh = o.nh
    for trial in range(4):
        before = random_state(o, lat, rng)
        plane = 1 if sign > 0 else nh[axis] - 2
        tang = [a for a in range(3) if a != axis]
        a_rng = (0, nh[tang[0]] - 1)
        b_rng = (0, nh[tang[1]] - 1) if dim == 3 else (0, 0)
        u = rng.uniform(-0.05, 0.05, 3)
        if dim == 2:
            u[2] = 0.0
            # 2D faces impose ux only: normal ux on x faces, tangential on y faces
            u[1] = 0.0
        rho_bc = rng.uniform(0.95, 1.05)
        o.zou_he_face(kind, plane, a_rng, b_rng, u=u, rho=rho_bc)
        after = o.f_halo()
        # every other plane is untouched
        mask = np.ones(nh, bool)
        idx = [slice(None)] * 3
        idx[axis] = plane
        mask[tuple(idx)] = False
        assert np.array_equal(after[mask], before[mask])
        b = face_nodes(before, axis, plane)
        a = face_nodes(after, axis, plane)
        if dens:
            ut = np.array([0.0, u[1], u[2]]) if dim == 3 else np.zeros(3)
            check_face_nodes(lat, b, a, axis, sign, rho_imp=rho_bc, ut=ut, dim=dim)
        else:
            rs = REFERENCE_DENSITY[kind](b, u) if kind in REFERENCE_DENSITY else None
            check_face_nodes(lat, b, a, axis, sign, u_imp=u, dim=dim, rho_star=rs)
            if kind in REFERENCE_DENSITY:
                # at u = 0 (the keyword's use) the face is plain Zou-He
                before = random_state(o, lat, rng)
                o.zou_he_face(kind, plane, a_rng, b_rng, u=(0.0, 0.0, 0.0))
                check_face_nodes(lat, face_nodes(before, axis, plane),
                                 face_nodes(o.f_halo(), axis, plane), axis, sign,
                                 u_imp=np.zeros(3), dim=dim)


# Face sets per keyword, src/fix_LBM_PSM_BC.cpp:274-400 (3D) and :223-271
# (2D), in application order: (kind, velocity factor of U, density face).
# e = 1; ranges are the interior [1, n-2] of the halo'd index, except the
# 2D xFlow y faces [2, n-3] (:239-242).
KEYWORDS_3D = {
    1: [(13, -1.0), (14, 1.0)],
    2: [(10, 1.0), (12, None), (13, 1.0), (14, 1.0), (15, 1.0), (16, 1.0)],
    3: [(13, 0.0), (14, 0.0)],
    4: [(10, 1.0), (12, None), (13, 0.0), (14, 0.0), (15, 0.0), (16, 0.0)],
    5: [(10, 0.0), (11, 0.0), (13, 0.0), (14, 0.0), (15, 0.0), (16, 0.0)],
}
KEYWORDS_2D = {
    1: [(3, -1.0), (4, 1.0)],
    2: [(0, 1.0), (2, None), (3, 1.0), (4, 1.0)],
    3: [(3, 0.0), (4, 0.0)],
    4: [(0, 1.0), (2, None), (3, 0.0), (4, 0.0)],
    5: [(0, 0.0), (1, 0.0), (3, 0.0), (4, 0.0)],
}


def _face_region(dim, kind, nh, keyword):
    """Boolean mask of the nodes one face of a keyword writes."""
    _, axis, sign, _ = FACES[kind]
    m = np.zeros(nh, bool)
    plane = 1 if sign > 0 else nh[axis] - 2
    rng = []
    for a in range(3):
        if a == axis:
            rng.append(slice(plane, plane + 1))
        elif dim == 2 and a == 2:
            rng.append(slice(0, 1))
        elif dim == 2 and keyword == 2 and axis == 1:
            rng.append(slice(2, nh[a] - 2))      # [2, n-3]
        else:
            rng.append(slice(1, nh[a] - 1))      # [1, n-2]
    m[tuple(rng)] = True
    return m


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("keyword", [1, 2, 3, 4, 5],
                         ids=["shear", "xFlow", "channel", "velIn-pressOut", "closedBox"])
def test_keyword_faces_and_overwrite_order(dim, keyword):
    lat = D3Q19 if dim == 3 else D2Q9
    U = 0.013
    u_unit = O.lattice_units(U)
    o = O.Oracle(dim, (7, 6, 5), (0, 0, 0), 0.7, unit=u_unit)
    rng = np.random.default_rng(7 * keyword + dim)
    before = random_state(o, lat, rng)
    O.LIB.lbo_apply_bc(o._h, keyword, U)
    after = o.f_halo()
    nh = o.nh
    faces = (KEYWORDS_3D if dim == 3 else KEYWORDS_2D)[keyword]
    regions = [_face_region(dim, k, nh, keyword) for k, _ in faces]
    touched = np.zeros(nh, bool)
    for r in regions:
        touched |= r
    assert np.array_equal(after[~touched], before[~touched])
    for idx, (kind, fac) in enumerate(faces):
        _, axis, sign, dens = FACES[kind]
        later = np.zeros(nh, bool)
        for r in regions[idx + 1:]:
            later |= r
        own = regions[idx] & ~later
        assert own.any()
        # no later face touched `own`: its final populations meet this
        # face's conditions (an earlier face at an edge may have changed
        # some of this face's known populations, so condition 1 is not
        # checked here)
        a = after[own]
        if dens:
            check_face_nodes(lat, None, a, axis, sign, rho_imp=1.0, dim=dim)
        else:
            u = np.array([fac * U, 0.0, 0.0])
            check_face_nodes(lat, None, a, axis, sign, u_imp=u, dim=dim)
