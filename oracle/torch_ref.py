"""Differentiable FP64 restatement of the Ah-grid SWM right-hand side and RK4 in PyTorch (CPU).

TEST INFRASTRUCTURE ONLY: the reference gradient for the device adjoint.  The formulas are
those of the C oracle (swm_oracle.c), written with tensor operations so that autograd gives
the exact transpose of the discrete forward map (up to rounding); the discrete operators
themselves (SBP matrices, halo map, sync weights, metric) are taken from the C oracle:
  - SBP derivative D (n x n): the oracle's sbp_apply_panel applied to unit vectors
    (sbp_operator_mod.f90:47-228);
  - depth-0 halo of each panel edge: the oracle's exchange map (exchange_factory_mod.f90:188-299);
  - scalar sync halo_Ah_scalar_sync_mod.f90:20-93, covariant vector sync
    ecs_Ah_vec_sync_mod.f90:34-125 (called with (gy, gx)), div_ah_sbp_mod.f90:24-98,
    grad_ah_sbp_mod.f90:26-90, curl_div_based_mod.f90:26-66, operator_swm_mod.f90:61-113,
    rk4_mod.f90:20-42.
Fields are (6, n, n) tensors indexed [panel, j, i] (mesh points only).  Orders 42, 43, 63.
"""
import numpy as np
import torch


class TorchSWM:
    def __init__(self, G, order):
        """G: oracle.Grid built with grad_order = div_kind = order and the physical constants."""
        assert order in (42, 43, 63)
        self.G, self.order, self.n = G, order, G.n
        n, hw = G.n, G.hw
        N = n - 1
        self.hx = (np.pi / 2) / N
        self.scale, self.grav = G.scale, G.grav
        # SBP matrix: column c of D = D applied to e_c (rows of a panel-0 field along x)
        D = np.zeros((n, n))
        for c in range(n):
            f = G.zeros()
            G.interior(f)[0, 0, c] = 1.0
            D[:, c] = G.interior(G.sbp(order, "x", f))[0, 0, :]
        self.D = torch.tensor(D, dtype=torch.float64)
        t = lambda a: torch.tensor(G.interior(a).copy(), dtype=torch.float64)  # noqa: E731
        self.J, self.Qi11, self.Qi12, self.Qi22, self.f = (t(G.geometry(k)) for k in ("J", "Qi11", "Qi12", "Qi22", "f"))
        w = G.sync_weights()
        self.qb, self.qt, self.ql, self.qr = (torch.tensor(w[k], dtype=torch.float64) for k in ("qb", "qt", "ql", "qr"))
        # depth-0 halo lines: source flat index into (6, n, n) and vector code, shape (6, n)
        src, vec = G.halo_map()
        ld = G.ld

        def to_interior(flat):
            q, rem = np.divmod(flat, ld * ld)
            jj, ii = np.divmod(rem, ld)
            return (q * n + (jj - hw)) * n + (ii - hw)

        rng = np.arange(n)
        lines = {"W": (rng + hw, np.full(n, hw - 1)), "E": (rng + hw, np.full(n, hw + n)),
                 "S": (np.full(n, hw - 1), rng + hw), "N": (np.full(n, hw + n), rng + hw)}
        self.hidx, self.hcode = {}, {}
        for key, (jj, ii) in lines.items():
            s = np.stack([src[p, jj, ii] for p in range(6)])
            self.hidx[key] = torch.tensor(to_interior(s), dtype=torch.long)
            self.hcode[key] = torch.tensor(np.stack([vec[p, jj, ii] for p in range(6)]), dtype=torch.long)

    # ---- halo and syncs ----
    def _halo(self, f, key):
        return f.reshape(-1)[self.hidx[key]]

    def _halo_vec(self, X, Y, key):
        a, b, c = self._halo(X, key), self._halo(Y, key), self.hcode[key]
        x = torch.where((c & 1) == 1, b, a)
        z = torch.where((c & 1) == 1, a, b)
        return torch.where((c & 2) == 2, -x, x), torch.where((c & 4) == 4, -z, z)

    def scalar_sync(self, f):
        n = self.n
        W, E, S, Nn = (self._halo(f, k) for k in ("W", "E", "S", "N"))
        o = f.clone()
        m = slice(1, n - 1)
        o[:, m, 0] = 0.5 * (W[:, m] + f[:, m, 0])
        o[:, m, n - 1] = 0.5 * (f[:, m, n - 1] + E[:, m])
        o[:, 0, m] = 0.5 * (f[:, 0, m] + S[:, m])
        o[:, n - 1, m] = 0.5 * (f[:, n - 1, m] + Nn[:, m])
        o[:, 0, 0] = (f[:, 0, 0] + W[:, 0] + S[:, 0]) / 3.0
        o[:, n - 1, 0] = (f[:, n - 1, 0] + W[:, n - 1] + Nn[:, 0]) / 3.0
        o[:, 0, n - 1] = (f[:, 0, n - 1] + E[:, 0] + S[:, n - 1]) / 3.0
        o[:, n - 1, n - 1] = (f[:, n - 1, n - 1] + E[:, n - 1] + Nn[:, n - 1]) / 3.0
        return o

    def vec_sync_covariant(self, gx, gy):
        """sync_grad_on_edges(U=gy, V=gx) after the vector exchange."""
        n = self.n
        hW, hE, hS, hN = (self._halo_vec(gx, gy, k) for k in ("W", "E", "S", "N"))
        U, V = gy, gx
        Uw, Vw = hW[1], hW[0]
        Ue, Ve = hE[1], hE[0]
        Us, Vs = hS[1], hS[0]
        Un, Vn = hN[1], hN[0]
        qb, qt, ql, qr = self.qb, self.qt, self.ql, self.qr
        Uo, Vo = U.clone(), V.clone()
        m = slice(1, n - 1)
        e = n - 1
        Uo[:, 0, 0] = (U[:, 0, 0] + Uw[:, 0] + Us[:, 0] + qb[0] * Vs[:, 0]) / 3.0
        Vo[:, 0, 0] = (V[:, 0, 0] + Vw[:, 0] + Vs[:, 0] + ql[0] * Uw[:, 0]) / 3.0
        Uo[:, 0, e] = (U[:, 0, e] + Ue[:, 0] + Us[:, e] + qb[e] * Vs[:, e]) / 3.0
        Vo[:, 0, e] = (V[:, 0, e] + Ve[:, 0] + Vs[:, e] + qr[0] * Ue[:, 0]) / 3.0
        Uo[:, e, 0] = (U[:, e, 0] + Uw[:, e] + Un[:, 0] + qt[0] * Vn[:, 0]) / 3.0
        Vo[:, e, 0] = (V[:, e, 0] + Vw[:, e] + Vn[:, 0] + ql[e] * Uw[:, e]) / 3.0
        Uo[:, e, e] = (U[:, e, e] + Ue[:, e] + Un[:, e] + qt[e] * Vn[:, e]) / 3.0
        Vo[:, e, e] = (V[:, e, e] + Ve[:, e] + Vn[:, e] + qr[e] * Ue[:, e]) / 3.0
        Uo[:, m, 0] = 0.5 * (Uw[:, m] + U[:, m, 0])
        Vo[:, m, 0] = 0.5 * (V[:, m, 0] + Vw[:, m] + ql[m] * Uw[:, m])
        Uo[:, m, e] = 0.5 * (U[:, m, e] + Ue[:, m])
        Vo[:, m, e] = 0.5 * (V[:, m, e] + Ve[:, m] + qr[m] * Ue[:, m])
        Vo[:, 0, m] = 0.5 * (Vs[:, m] + V[:, 0, m])
        Uo[:, 0, m] = 0.5 * (Us[:, m] + U[:, 0, m] + qb[m] * Vs[:, m])
        Vo[:, e, m] = 0.5 * (V[:, e, m] + Vn[:, m])
        Uo[:, e, m] = 0.5 * (U[:, e, m] + Un[:, m] + qt[m] * Vn[:, m])
        return Vo, Uo  # (gx, gy)

    # ---- operators ----
    def dx(self, f):
        return torch.matmul(f, self.D.T)

    def dy(self, f):
        return torch.matmul(self.D, f)

    def div(self, u, v):
        J = self.J
        return self.scalar_sync((self.dx(J * u) + self.dy(J * v)) / (J * self.hx * self.scale))

    def grad(self, f):
        return self.vec_sync_covariant(self.dx(f) / (self.hx * self.scale), self.dy(f) / (self.hx * self.scale))

    def curl(self, u, v):
        return self.div(v / self.J, -u / self.J)

    def rhs(self, h, u, v):
        grav = self.grav
        ut = self.Qi11 * u + self.Qi12 * v
        vt = self.Qi12 * u + self.Qi22 * v
        hu, hv = h * ut, h * vt
        K = (u * ut + v * vt) / 2.0 + grav * h
        gx, gy = self.grad(K)
        z = self.f + self.curl(u, v)
        du = -gx + z * hv / h * self.J
        dv = -gy - z * hu / h * self.J
        return -self.div(hu, hv), du, dv

    def rk4(self, y, dt, nsteps):
        c1, c2, c3 = 0.5 * dt, dt / 6.0, dt / 3.0
        for _ in range(nsteps):
            k1 = self.rhs(*y)
            k2 = self.rhs(*(a + c1 * b for a, b in zip(y, k1)))
            k3 = self.rhs(*(a + c1 * b for a, b in zip(y, k2)))
            k4 = self.rhs(*(a + dt * b for a, b in zip(y, k3)))
            y = tuple(a + c2 * b1 + c3 * b2 + c3 * b3 + c2 * b4 for a, b1, b2, b3, b4 in zip(y, k1, k2, k3, k4))
        return y
