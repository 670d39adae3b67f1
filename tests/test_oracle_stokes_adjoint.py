"""Pins the oracle's viscosity-gradient formula (orc_stokes_grad_eta) on the
CPU: forward and adjoint steady PT Stokes solves with the oracle's own
update_stress!/update_velocity!/bc! (residual < 1e-12), then dJ/dlog(eta)
from the formula against central differences of J computed by more oracle
solves.  A sign, factor or edge-weight mistake in the formula fails here."""
import numpy as np
import pytest

import oracle as O


def _locs(nd):
    C, V = O.CENTER, O.VERTEX
    if nd == 2:
        return [(V, C), (C, V)], [(C, C), (C, C), (V, V)]
    return [(V, C, C), (C, V, C), (C, C, V)], [(C, C, C)] * 3 + [(V, V, C), (V, C, V), (C, V, V)]


class Problem:
    def __init__(self, dims):
        self.nd = nd = len(dims)
        self.g = O.Grid((-0.5,) * nd, (1.0,) * nd, dims)
        vl, tl = _locs(nd)
        self.vl, self.tl = vl, tl
        mk = lambda ls: [O.HField(self.g, l) for l in ls]
        self.V, self.rV, self.fV = mk(vl), mk(vl), mk(vl)
        self.tau, self.told = mk(tl), mk(tl)
        self.Pr, self.dV, self.eta, self.eml = (O.HField(self.g, O.CENTER) for _ in range(4))
        self.p = O.pt_params(self.g)
        self.G = 1.0
        self.dt = min(self.g.spacing) ** 2 / 1e-4 / nd / 2.1

    def field(self, sigma, seed, corr=3):
        rng = np.random.default_rng(seed)
        a = rng.standard_normal(self.eta.arr.shape)
        k = np.ones(corr) / corr
        for ax in range(a.ndim):
            a = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, a)
        return sigma * (a - a.mean()) / a.std()

    def set_logeta(self, la):
        self.eta.arr[...] = np.exp(la)
        O.maxloc(self.g, self.eml, self.eta)

    def spec(self, normal):
        s = []
        for D in range(self.nd):
            d = {q: O.Neumann() for q in range(self.nd)}
            a, b = normal.get(D, (0.0, 0.0))
            d[D] = (O.Dirichlet(a), O.Dirichlet(b))
            s.append((self.V[D], d))
        return s

    def solve(self, normal, forcing=False, eps=1e-12, ncheck=200, max_iter=400000):
        for f in self.V + self.tau + [self.Pr]:
            f.arr[...] = 0.0
        spec = self.spec(normal)
        O.bc_apply(self.g, spec)
        rspec = [(self.rV[D], {D: O.Dirichlet()}) for D in range(self.nd)]
        it = 0
        while it < max_iter:
            for _ in range(ncheck):
                O.update_stress(self.g, self.tau, self.Pr, self.dV, self.V, self.told, self.eta, 0.0, self.G,
                                self.dt, self.p["dtau_Pr"], self.p["dtau_r"])
                O.update_velocity(self.g, self.V, self.rV, self.Pr, self.tau, None, 0.0, self.p["nudtau"],
                                  eta_ml=self.eml, G=self.G, dt=self.dt, fV=self.fV if forcing else None)
                O.bc_apply(self.g, spec)
            it += ncheck
            O.bc_apply(self.g, rspec)
            err = [O.maxabs(self.dV)] + [O.maxabs(r) for r in self.rV]
            assert all(np.isfinite(err))
            if max(err) < eps:
                return it
        raise AssertionError(f"no convergence: {err}")

    def top(self, f):
        # last-dimension index n (storage index n + 1), interior in the others
        sl = (f.n[-1] + 1,) + tuple(slice(2, 2 + m) for m in reversed(f.n[:-1]))
        return f.arr[sl]


NORMAL = {0: (1.0, -1.0)}


@pytest.mark.parametrize("dims", [(9, 7), (6, 5, 4)])
def test_oracle_grad_formula_matches_central_differences(dims):
    P = Problem(dims)
    nd = P.nd
    normal = dict(NORMAL)
    normal[nd - 1] = (-1.0, 1.0)
    la0 = P.field(0.5, seed=1)
    P.set_logeta(la0 + P.field(0.3, seed=7))
    P.solve(normal)
    Vobs = [P.top(P.V[D]).copy() for D in range(nd - 1)]

    def J_at(la):
        P.set_logeta(la)
        P.solve(normal)
        return sum(0.5 * float(np.sum((P.top(P.V[D]) - Vobs[D]) ** 2)) for D in range(nd - 1))

    J = J_at(la0)
    tau_f = [O.HField(P.g, l) for l in P.tl]
    for a, b in zip(tau_f, P.tau):
        a.arr[...] = b.arr
    for f in P.fV:
        f.arr[...] = 0.0
    for D in range(nd - 1):
        P.top(P.fV[D])[...] = -(P.top(P.V[D]) - Vobs[D])
    P.solve({}, forcing=True)
    grad = O.HField(P.g, O.CENTER)
    O.stokes_grad_eta(P.g, tau_f, P.tau, P.eta, grad, logscale=True)
    rng = np.random.default_rng(5)
    h = 1e-4
    for _ in range(2):
        v = rng.standard_normal(la0.shape)
        fd = (J_at(la0 + h * v) - J_at(la0 - h * v)) / (2 * h)
        ad = float(np.sum(grad.arr * v))
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    assert J > 0
