"""FP64 PyTorch-CPU restatement of the plain forward step on a periodic box
(test infrastructure: the autograd reference of the discrete adjoint).

One RK3 step of time_integration_rk (src/mainloop.F90:392-476) with the
stage order of the main path, on haloed arrays (nodes -hm..n+hm):
filterq (src/comsolver.F90:514-632, periodic wrap of the interior), qswap
(src/parallel.F90:4849-5319: halo copy, duplicate average and q2fvar of the
strips, axis by axis), gradcal, sigma/heat flux and the axis fluxes
G = Fv - F (convrsdcal6 + diffrsdcal6, viscous halos from the periodic
image), their derivatives, the RK3 combine and updatefvar.  The interior
primitives stay stale across the filter, as in the reference.  The line
operators are the checker's df_compact and compact_filter (bit-identical to
the reference modules) applied as dense matrices.  The objective is the
enstrophy statistic (enstophycal, src/statistic.F90:870-925) of the final
state after a qswap.
"""
import numpy as np
import torch

import pyoracle as po

HM = 5


def line_matrix(op, n):
    w = n + 1 + 2 * HM
    M = np.zeros((n + 1, w))
    for c in range(w):
        e = np.zeros(w)
        e[c] = 1.0
        M[:, c] = po.df_compact(3, e) if op == "d" else po.compact_filter(3, e)
    return torch.tensor(M, dtype=torch.float64)


class Box:
    def __init__(self, n, reynolds=1600.0, mach=0.1, prandtl=0.72, gamma=1.4, ref_tem=273.15,
                 sutherland=110.3, lengths=(2 * np.pi,) * 3):
        self.n = tuple(n)
        self.Md = [line_matrix("d", k) for k in self.n]
        self.Mf = [line_matrix("f", k) for k in self.n]
        g, ma = gamma, mach
        self.const1 = 1.0 / (g * (g - 1.0) * (ma * ma))
        self.const2 = g * (ma * ma)
        self.const5 = (g - 1.0) * (ma * ma)
        self.const6 = 1.0 / (g - 1.0)
        self.reynolds, self.prandtl = reynolds, prandtl
        self.tempconst = sutherland / ref_tem
        self.tempconst1 = 1.0 + self.tempconst
        self.h = [lengths[d] / self.n[d] for d in range(3)]
        self.xi = [1.0 / h for h in self.h]
        self.jac = self.h[0] * self.h[1] * self.h[2]
        self.I = tuple(slice(HM, HM + k + 1) for k in self.n)

    # -------------------------------------------------- pointwise physics
    def q2fvar(self, q):
        rho = q[0]
        u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
        p = (q[4] - 0.5 * rho * (u * u + v * v + w * w)) / self.const6
        return torch.stack([rho, u, v, w, p, p / rho * self.const2])

    def miucal(self, t):
        return t * torch.sqrt(t) * self.tempconst1 / (t + self.tempconst)

    def initial(self, u0):
        """State from the velocity at the unique nodes (3, n0, n1, n2):
        rho = 1, p = 1/const2, T = 1, q by fvar2q's temperature form on the
        interior (node n copies node 0), halos zero."""
        n = self.n
        idx = [torch.arange(k + 1) % k for k in n]
        u = u0[:, idx[0]][:, :, idx[1]][:, :, :, idx[2]]
        shape = tuple(k + 1 + 2 * HM for k in n)
        rho = torch.ones_like(u[0])
        prs = torch.full_like(u[0], 1.0 / self.const2)
        tmp = prs / rho * self.const2
        P = torch.zeros((6,) + shape, dtype=torch.float64)
        Q = torch.zeros((5,) + shape, dtype=torch.float64)
        Pi = torch.stack([rho, u[0], u[1], u[2], prs, tmp])
        Qi = torch.stack([rho, rho * u[0], rho * u[1], rho * u[2],
                          rho * (tmp * self.const1 + 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2))])
        P = P.clone()
        Q = Q.clone()
        P[(slice(None),) + self.I] = Pi
        Q[(slice(None),) + self.I] = Qi
        return Q, P

    # ----------------------------------------------------- line operators
    def _line_nodes(self, d, wrap):
        n = self.n[d]
        t = torch.arange(-HM, n + HM + 1)
        if wrap:
            t = torch.where(t < 0, t + n, torch.where(t > n, t - n, t))
        return t + HM

    def _apply(self, M, X, axis):
        return torch.movedim(torch.movedim(X, axis, -1) @ M.T, -1, axis)

    def _lines(self, F, d, wrap):
        """F (c, haloed) -> lines along d over -hm..n+hm (wrapped from the
        interior, or the halos as stored), interior on the other axes."""
        sl = [slice(None)] + [self.I[e] for e in range(3)]
        sl[1 + d] = slice(None)
        X = F[tuple(sl)]
        return torch.index_select(X, 1 + d, self._line_nodes(d, wrap)) if wrap else X

    def filterq(self, Q):
        for d in range(3):
            out = self._apply(self.Mf[d], self._lines(Q, d, True), 1 + d)
            Q = Q.clone()
            Q[(slice(None),) + self.I] = out
        return Q

    def qswap(self, Q, P):
        for d in range(3):
            n = self.n[d]
            Q = Q.clone()
            P = P.clone()

            def at(t):
                sl = [slice(None)] + [self.I[e] for e in range(3)]
                sl[1 + d] = t
                return tuple(sl)

            Q[at(slice(0, HM))] = Q[at(slice(HM + n - HM, HM + n))]            # -hm..-1 <- n-hm..n-1
            Q[at(slice(HM + n + 1, HM + n + 1 + HM))] = Q[at(slice(HM + 1, HM + 1 + HM))]
            avg = 0.5 * (Q[at(slice(HM, HM + 1))] + Q[at(slice(HM + n, HM + n + 1))])
            Q[at(slice(HM, HM + 1))] = avg
            Q[at(slice(HM + n, HM + n + 1))] = avg
            for sl in (slice(HM + n, HM + n + 1 + HM), slice(0, HM + 1)):
                P[at(sl)] = self.q2fvar(Q[at(sl)])
        return Q, P

    def gradcal(self, P):
        dvel = [[None] * 3 for _ in range(3)]
        dtmp = [None] * 3
        for d in range(3):
            lines = self._lines(P, d, False)
            D = self._apply(self.Md[d], lines, 1 + d) * self.xi[d]
            for m in range(3):
                dvel[m][d] = D[1 + m]
            dtmp[d] = D[5]
        return dvel, dtmp

    def sigma(self, dvel, dtmp, P):
        miu = self.miucal(P[5]) / self.reynolds
        d = dvel
        s11, s22, s33 = d[0][0], d[1][1], d[2][2]
        s12, s13, s23 = 0.5 * (d[0][1] + d[1][0]), 0.5 * (d[0][2] + d[2][0]), 0.5 * (d[1][2] + d[2][1])
        skk = (1.0 / 3.0) * (s11 + s22 + s33)
        miu2 = 2.0 * miu
        hcc = (miu / self.prandtl) / self.const5
        g = [miu2 * (s11 - skk), miu2 * s12, miu2 * s13, miu2 * (s22 - skk), miu2 * s23,
             miu2 * (s33 - skk)]
        u, v, w = P[1], P[2], P[3]
        qf = [hcc * dtmp[0] + g[0] * u + g[1] * v + g[2] * w,
              hcc * dtmp[1] + g[1] * u + g[3] * v + g[4] * w,
              hcc * dtmp[2] + g[2] * u + g[4] * v + g[5] * w]
        return g, qf

    def gflux(self, dvel, dtmp, Q, P):
        """G_d (5, lines along d over -hm..n+hm, interior on the others)."""
        Pi = P[(slice(None),) + self.I]
        g, qf = self.sigma(dvel, dtmp, Pi)
        sidx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
        out = []
        for d in range(3):
            Ql = self._lines(Q, d, False)
            Pl = self._lines(P, d, False)
            # viscous part on the line: interior values, the periodic image
            # on the halos
            wi = self._line_nodes(d, True) - HM
            sv = [torch.index_select(g[sidx[d][c]], d, wi) for c in range(3)]
            qv = torch.index_select(qf[d], d, wi)
            jc, xn = self.jac, self.xi[d]
            uu = xn * Pl[1 + d]
            p = Pl[4]
            G = []
            for m in range(5):
                qm = Ql[m]
                if m == 0:
                    fi = jc * qm * uu
                elif m == 4:
                    fi = jc * (qm + p) * uu
                elif m == d + 1:
                    fi = jc * (qm * uu + xn * p)
                else:
                    fi = jc * (qm * uu)
                if m == 0:
                    G.append(-fi)
                else:
                    sv_m = sv[m - 1] if m < 4 else qv
                    G.append((sv_m * xn) * jc - fi)
            out.append(torch.stack(G))
        return out

    def stage(self, Q, P, qsave, s, dt):
        rk = [(1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)][s - 1]
        Q = self.filterq(Q)
        Q, P = self.qswap(Q, P)
        dvel, dtmp = self.gradcal(P)
        G = self.gflux(dvel, dtmp, Q, P)
        qrhs = sum(self._apply(self.Md[d], G[d], 1 + d) for d in range(3))
        Qi = Q[(slice(None),) + self.I]
        if s == 1:
            qsave = Qi * self.jac
        qn = (rk[0] * qsave + rk[1] * Qi * self.jac + rk[2] * qrhs * dt) / self.jac
        Q = Q.clone()
        P = P.clone()
        Q[(slice(None),) + self.I] = qn
        P[(slice(None),) + self.I] = self.q2fvar(qn)
        return Q, P, qsave

    def step(self, Q, P, dt):
        qsave = None
        for s in (1, 2, 3):
            Q, P, qsave = self.stage(Q, P, qsave, s, dt)
        return Q, P

    def enstrophy(self, Q, P):
        Q, P = self.qswap(Q, P)
        dvel, _ = self.gradcal(P)
        sl = tuple(slice(1, None) for _ in range(3))
        d = [[dvel[m][e][sl] for e in range(3)] for m in range(3)]
        o1 = d[2][1] - d[1][2]
        o2 = d[0][2] - d[2][0]
        o3 = d[1][0] - d[0][1]
        rho = P[0][self.I][sl]
        n3 = float(self.n[0] * self.n[1] * self.n[2])
        return 0.5 * torch.sum(rho * (o1 * o1 + o2 * o2 + o3 * o3)) / n3

    def objective_state(self, Q, P, nsteps, dt):
        for _ in range(nsteps):
            Q, P = self.step(Q, P, dt)
        return self.enstrophy(Q, P)

    def objective(self, u0, nsteps, dt):
        Q, P = self.initial(u0)
        return self.objective_state(Q, P, nsteps, dt)


def to_host(Q, P):
    """Torch (field, i, j, k) state -> the haloed host arrays of
    AstrSolver.upload_state."""
    q = np.asfortranarray(np.moveaxis(Q.detach().numpy(), 0, -1))
    pr = P.detach().numpy()
    return {"q": q, "rho": np.asfortranarray(pr[0]),
            "vel": np.asfortranarray(np.moveaxis(pr[1:4], 0, -1)),
            "prs": np.asfortranarray(pr[4]), "tmp": np.asfortranarray(pr[5])}
