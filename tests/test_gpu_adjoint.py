"""The library's discrete adjoint against central finite differences of the
FP64 forward run (the same library, same inputs).

The adjoint is exact for the discrete forward model, so the finite
differences agree with it up to their own truncation and rounding error.
Central differences with h = 1e-6 carry an h^2 truncation error that is
~1e-5 relative for the body-force components of the 96^3 case (the loss
is strongly nonlinear in F over 500 steps; halving h divides the gap by 4).
The finite differences are therefore Richardson-extrapolated from h and
h/2, (4 D(h/2) - D(h)) / 3, which removes the h^2 term.  The bar is 1e-6
relative per component, taken against max(|fd|, 1e-3 x the largest checked
gradient) so that components near zero are compared at the scale of the
others.
"""
import numpy as np
import pytest
import torch

import lbmpsm
from lbmpsm import harness as H
from lbmpsm.autograd import HydroRun, RunSpec, run_losses
from scenarios import lattice_box

pytestmark = pytest.mark.gpu

RTOL = 1e-6
H_FD = 1e-6


def _owner_sum_t(X, owner, nl):
    return torch.zeros((nl, X.shape[1]), dtype=X.dtype).index_add(0, owner, X)


def _owner_sum_np(X, owner, nl):
    out = np.zeros((nl, X.shape[1]))
    np.add.at(out, owner, X)
    return out


class Problem:
    """Owner-level parameters -> list-level inputs -> loss."""

    def __init__(self, spec, lst, weights):
        self.spec = spec
        self.owner = np.asarray(lst.owner, dtype=np.int64)
        self.nl = lst.nlocal
        self.wF, self.wT, self.wS = weights

    def loss_t(self, F, T, S):
        own = torch.from_numpy(self.owner)
        Fo = _owner_sum_t(F, own, self.nl)
        To = _owner_sum_t(T, own, self.nl)
        return (self.wF * (Fo[:, 0] ** 2).sum() + self.wT * (Fo[:, 1] * To[:, 2]).sum()
                + self.wS * S[:, 3].sum())

    def loss_np(self, F, T, S):
        Fo = _owner_sum_np(F, self.owner, self.nl)
        To = _owner_sum_np(T, self.owner, self.nl)
        return (self.wF * np.sum(Fo[:, 0] ** 2) + self.wT * np.sum(Fo[:, 1] * To[:, 2])
                + self.wS * np.sum(S[:, 3]))

    def adjoint(self, F_lbm, v, w, tau):
        tF = torch.tensor(F_lbm, dtype=torch.float64, requires_grad=True)
        tv = torch.tensor(v, dtype=torch.float64, requires_grad=True)
        tw = torch.tensor(w, dtype=torch.float64, requires_grad=True)
        tt = torch.tensor(tau, dtype=torch.float64, requires_grad=True)
        own = torch.from_numpy(self.owner)
        F, T, S = HydroRun.apply(tF, tv[own], tw[own], tt, self.spec)
        L = self.loss_t(F, T, S)
        L.backward()
        return L.detach().item(), np.concatenate([tF.grad.numpy(), tv.grad.numpy().ravel(),
                                         tw.grad.numpy().ravel(), [tt.grad.item()]])

    def loss_at(self, theta, nl):
        F_lbm = theta[:3]
        v = theta[3:3 + 3 * nl].reshape(nl, 3)
        w = theta[3 + 3 * nl:3 + 6 * nl].reshape(nl, 3)
        tau = theta[-1]
        return self.loss_np(*run_losses(F_lbm, v[self.owner], w[self.owner], tau, self.spec))


def check(prob, F_lbm, v, w, tau, comps, label):
    nl = prob.nl
    L0, g = prob.adjoint(F_lbm, v, w, tau)
    theta = np.concatenate([F_lbm, v.ravel(), w.ravel(), [tau]])
    fd = np.zeros(len(comps))

    def central(k, h):
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        return (prob.loss_at(tp, nl) - prob.loss_at(tm, nl)) / (2.0 * h)

    for c, k in enumerate(comps):
        fd[c] = (4.0 * central(k, 0.5 * H_FD) - central(k, H_FD)) / 3.0
    ga = g[comps]
    floor = 1e-3 * np.max(np.abs(fd))
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), floor)
    print(f"\n  {label}: loss {L0:.6e}, max |grad| {np.max(np.abs(fd)):.3e}, "
          f"max rel err {rel.max():.2e}", flush=True)
    for k, a, b, r in zip(comps, ga, fd, rel):
        print(f"    comp {k:4d}: adjoint {a: .10e}  fd {b: .10e}  rel {r:.1e}", flush=True)
    assert rel.max() <= RTOL
    return rel.max()


def test_adjoint_2d_disks_with_ghost_and_body_force():
    print("\n[case] adjoint D2Q9 40x32, 2 disks (one with a periodic image), 30 steps, "
          "checkpoints every 7", flush=True)
    box = lattice_box(2, (40, 32, 1), (1, 1, 1))
    atoms = H.Atoms(tag=np.array([4, 9], np.int32),
                    x=np.array([[1.6, 15.3, 0.0], [24.7, 9.4, 0.0]]),
                    v=np.zeros((2, 3)), omega=np.zeros((2, 3)), radius=np.array([4.5, 5.2]))
    lst = H.build_list(atoms, box, cutoff=7.0)
    assert lst.n == 3
    spec = RunSpec(dim=2, n=(40, 32), tag=lst.tag, x=lst.x, radius=lst.radius, steps=30,
                   checkpoint_every=7)
    prob = Problem(spec, lst, (1.0, 0.7, 0.3))
    v = np.array([[2e-3, -1e-3, 0.0], [-1.5e-3, 2.5e-3, 0.0]])
    w = np.array([[0.0, 0.0, 3e-4], [0.0, 0.0, -2e-4]])
    F_lbm = np.array([2e-5, -1e-5, 0.0])
    nl = 2
    comps = [0, 1, 3, 4, 6, 7, 3 + 3 * nl + 2, 3 + 3 * nl + 5, 3 + 6 * nl]
    check(prob, F_lbm, v, w, 0.8, comps, "2D")


def test_adjoint_rejects_unsupported_contexts():
    cfg = lbmpsm.lattice_units_config(3, (16, 16, 16), (1, 1, 1), 0.7, storage="f32")
    lat = lbmpsm.Lattice(cfg)
    one = dict(tag=np.array([1], np.int32), x=np.array([[8.0, 8.0, 8.0]]), v=np.zeros((1, 3)),
               omega=np.zeros((1, 3)), radius=np.array([3.0]))
    with pytest.raises(lbmpsm.LbmPsmError):
        lat.adjoint_forward(5, 2, **one)
    lat.close()
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, (16, 16, 16), (1, 0, 1), 0.7))
    lat.set_bc("shear")
    with pytest.raises(lbmpsm.LbmPsmError):
        lat.adjoint_forward(5, 2, **one)
    lat.close()
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, (16, 16, 16), (1, 1, 1), 0.7))
    lat._n = 1
    with pytest.raises(lbmpsm.LbmPsmError):
        lat.adjoint_backward(np.zeros((1, 3)))
    lat.close()


def test_adjoint_27_spheres_96cube():
    """96^3 periodic, 27 spheres d = 12 on a 3x3x3 grid with seeded +-3
    jitter, T = 500 steps, checkpoints every 50, loss = sum over spheres of
    the squared x-drag; 20 seeded components of (F_lbm, v, omega, tau)."""
    print("\n[case] adjoint D3Q19 96^3, 27 spheres, 500 steps, checkpoints every 50", flush=True)
    rng = np.random.default_rng(404)
    g = np.array([16.0, 48.0, 80.0])
    pos = np.array([[a, b, c] for a in g for b in g for c in g]) + rng.uniform(-3, 3, (27, 3))
    box = lattice_box(3, (96, 96, 96), (1, 1, 1))
    atoms = H.Atoms(tag=np.arange(1, 28, dtype=np.int32), x=pos, v=np.zeros((27, 3)),
                    omega=np.zeros((27, 3)), radius=np.full(27, 6.0))
    lst = H.build_list(atoms, box, cutoff=8.0)
    spec = RunSpec(dim=3, n=(96, 96, 96), tag=lst.tag, x=lst.x, radius=lst.radius, steps=500,
                   checkpoint_every=50)
    prob = Problem(spec, lst, (1.0, 0.0, 0.0))
    v = rng.uniform(-1e-3, 1e-3, (27, 3))
    w = rng.uniform(-1e-4, 1e-4, (27, 3))
    F_lbm = np.array([5e-6, 0.0, 0.0])
    ntheta = 3 + 6 * 27 + 1
    comps = sorted(rng.choice(ntheta, size=20, replace=False).tolist())
    check(prob, F_lbm, v, w, 0.7, comps, "3D cfg")


def test_backward_leaves_the_forward_state():
    """The recomputation inside backward runs collides of its own; the slot
    forces, slots and populations the forward left stay as they were."""
    print("\n[case] adjoint leaves forward state, D3Q19 24^3, 2 spheres, 12 steps", flush=True)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, (24, 24, 24), (1, 1, 1), 0.75,
                                                     F_lbm=(2e-5, 0.0, 0.0)))
    one = dict(tag=np.array([1, 2], np.int32), x=np.array([[8.3, 11.6, 12.2], [15.1, 12.4, 11.0]]),
               v=np.array([[1e-3, 0.0, 0.0], [0.0, -1e-3, 0.0]]), omega=np.zeros((2, 3)),
               radius=np.array([3.5, 3.2]))
    lat.adjoint_forward(12, 5, **one)
    t0, sf0, us0, F0 = lat.slots()
    f0 = lat.populations()
    lat.adjoint_backward(np.ones((2, 3)))
    t1, sf1, us1, F1 = lat.slots()
    assert np.any(F0 != 0.0)
    assert np.array_equal(t0, t1) and np.array_equal(sf0, sf1) and np.array_equal(us0, us1)
    assert np.array_equal(F0, F1), "backward changed the slot forces"
    assert np.array_equal(f0, lat.populations())
    lat.close()
