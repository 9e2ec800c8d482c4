"""Mock of the LAMMPS side of `fix lbm-psm` (no LAMMPS is available here).

It supplies what the fix receives from LAMMPS and does with its results:

* the atom list in LAMMPS order: locals, then periodic ghost images built
  dimension by dimension like Comm::borders on one processor (each
  dimension images the atoms present when it starts, lower face first);
* `fix move linear` prescribed motion with periodic remap;
* the post_force bookkeeping of src/fix_LBM_PSM.cpp:223-263: f += F for
  locals, f = F for ghosts, then the reverse communication that sums ghost
  forces into their owners, and the v_tally stresslet sum;
* seeded random sequential addition (RSA) of non-overlapping spheres.

Ghost ordering, reverse-communication order and v_tally semantics are
LAMMPS internals that no reference test pins; this module defines them and
feeds the same lists to the GPU library and to the CPU oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Box:
    lo: np.ndarray          # boxlo
    prd: np.ndarray         # box lengths
    periodic: tuple         # per axis
    dim: int = 3

    @property
    def hi(self):
        return self.lo + self.prd


@dataclass
class Atoms:
    tag: np.ndarray         # int32 [n]
    x: np.ndarray           # [n,3]
    v: np.ndarray           # [n,3]
    omega: np.ndarray       # [n,3]
    radius: np.ndarray      # [n]
    x0: np.ndarray = field(default=None)  # original positions for fix move

    @property
    def n(self):
        return self.tag.shape[0]

    def copy(self):
        return Atoms(self.tag.copy(), self.x.copy(), self.v.copy(), self.omega.copy(),
                     self.radius.copy(), None if self.x0 is None else self.x0.copy())


@dataclass
class AtomList:
    """Locals followed by ghosts, as the fix sees atom->x etc."""
    tag: np.ndarray
    x: np.ndarray
    v: np.ndarray
    omega: np.ndarray
    radius: np.ndarray
    owner: np.ndarray       # local index each entry belongs to
    nlocal: int

    @property
    def n(self):
        return self.tag.shape[0]


def build_list(atoms: Atoms, box: Box, cutoff: float) -> AtomList:
    """Locals + periodic ghost images within `cutoff` of each periodic face."""
    tag = [atoms.tag]
    x = [atoms.x]
    v = [atoms.v]
    om = [atoms.omega]
    rad = [atoms.radius]
    owner = [np.arange(atoms.n)]
    cur_x = atoms.x.copy()
    cur = dict(tag=atoms.tag, v=atoms.v, omega=atoms.omega, radius=atoms.radius,
               owner=np.arange(atoms.n))
    for d in range(box.dim):
        if not box.periodic[d]:
            continue
        lo, hi, L = box.lo[d], box.hi[d], box.prd[d]
        base_x = cur_x
        base = dict(cur)
        add_x, add = [], {k: [] for k in base}
        for side in (0, 1):
            if side == 0:
                sel = base_x[:, d] < lo + cutoff
                shift = L
            else:
                sel = base_x[:, d] >= hi - cutoff
                shift = -L
            gx = base_x[sel].copy()
            gx[:, d] = gx[:, d] + shift
            add_x.append(gx)
            for k in base:
                add[k].append(base[k][sel])
        new_x = np.concatenate(add_x) if add_x else np.zeros((0, 3))
        x.append(new_x)
        tag.append(np.concatenate(add["tag"]))
        v.append(np.concatenate(add["v"]))
        om.append(np.concatenate(add["omega"]))
        rad.append(np.concatenate(add["radius"]))
        owner.append(np.concatenate(add["owner"]))
        cur_x = np.concatenate([base_x, new_x])
        for k in base:
            cur[k] = np.concatenate([base[k]] + add[k])
    return AtomList(
        tag=np.ascontiguousarray(np.concatenate(tag), dtype=np.int32),
        x=np.ascontiguousarray(np.concatenate(x), dtype=np.float64),
        v=np.ascontiguousarray(np.concatenate(v), dtype=np.float64),
        omega=np.ascontiguousarray(np.concatenate(om), dtype=np.float64),
        radius=np.ascontiguousarray(np.concatenate(rad), dtype=np.float64),
        owner=np.concatenate(owner),
        nlocal=atoms.n,
    )


def remap(x: np.ndarray, box: Box) -> np.ndarray:
    """Domain::remap for periodic axes (positions back into [lo, hi))."""
    y = x.copy()
    for d in range(box.dim):
        if box.periodic[d]:
            lo, L = box.lo[d], box.prd[d]
            y[:, d] = lo + np.mod(y[:, d] - lo, L)
    return y


def move_linear(atoms: Atoms, box: Box, step: int, dt: float) -> None:
    """fix move linear: x = x0 + v * (step * dt), then remap."""
    if atoms.x0 is None:
        atoms.x0 = atoms.x.copy()
    atoms.x = remap(atoms.x0 + atoms.v * (step * dt), box)


def reduce_to_owners(lst: AtomList, F, T):
    """post_force + reverse_comm: owners receive their own F/T plus every
    ghost image's, summed in list order."""
    Fo = np.zeros((lst.nlocal, 3))
    To = np.zeros((lst.nlocal, 3))
    for i in range(lst.n):
        o = lst.owner[i]
        Fo[o] += F[i]
        To[o] += T[i]
    return Fo, To


def rsa_spheres(box: Box, radii: np.ndarray, seed: int, gap: float = 0.0,
                region=None, max_tries: int = 2_000_000) -> np.ndarray:
    """Random sequential addition of spheres (largest first is *not* applied:
    radii are placed in the given order).  Periodic axes use the minimum
    image; on wall axes the sphere stays `gap` inside the box.  `region`
    optionally restricts centres to [lo, hi) per axis."""
    rng = np.random.default_rng(seed)
    n = radii.shape[0]
    pos = np.zeros((n, 3))
    placed = 0
    lo = box.lo.copy()
    hi = box.hi.copy()
    if region is not None:
        lo = np.maximum(lo, np.asarray(region[0], dtype=float))
        hi = np.minimum(hi, np.asarray(region[1], dtype=float))
    tries = 0
    while placed < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"RSA jammed after placing {placed} of {n} spheres")
        r = radii[placed]
        c = np.zeros(3)
        ok = True
        for d in range(3):
            if d >= box.dim:
                c[d] = 0.0
                continue
            if box.periodic[d]:
                c[d] = rng.uniform(lo[d], hi[d])
            else:
                a, b = lo[d] + r + gap, hi[d] - r - gap
                if b <= a:
                    ok = False
                    break
                c[d] = rng.uniform(a, b)
        if not ok:
            raise RuntimeError("sphere does not fit between the walls")
        if placed:
            dvec = pos[:placed] - c
            for d in range(box.dim):
                if box.periodic[d]:
                    L = box.prd[d]
                    dvec[:, d] -= L * np.round(dvec[:, d] / L)
            dist2 = np.sum(dvec[:, : box.dim] ** 2, axis=1)
            lim = (radii[:placed] + r + gap) ** 2
            if np.any(dist2 < lim):
                continue
        pos[placed] = c
        placed += 1
    return pos
