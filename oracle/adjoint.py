"""Discrete adjoint of the converged steady solution, on the CPU (numpy +
scipy): the checker of the device adjoint.

TEST INFRASTRUCTURE ONLY: used by tests/ (and the CPU leg of bench.py) as the
checker; the product never imports it.

The reference solver's converged state x* is a fixed point of its V-cycle,
and a fixed point of the V-cycle is a zero of the level-0 steady residual
R(x; theta) (the under-relaxation of umom/vmom cancels there, and the FAS
terms of the coarse levels vanish when the fine residual does):

  R_u(i,j)  = sum_k a_k (u_P - u_k) - areax (p_P - p_E)
              the unrelaxed u momentum of coefu + umom
              (Version-6-26-04.c:741-767, 815-860): a_w = h+(cxm, D),
              a_e = h-(cxp, D), a_s = h+(cym, D), a_n = h-(cyp, D) with
              h+-(c, D) = fmax(fabs(c), D) +- c, cxm = 0.5 (cx_P + cx_W),
              cx = 0.5 u areax, and the wall diffusion added to a_s on the
              first row and a_n on the last (V6:849-858);
  R_v(i,j)  the same for v (coefv + vmom, V6:769-793, 966-1016), with the
              side-wall adds amu * areay / dy on the first and last column;
  R_c(i,j)  = areax (u_P - u_W) + areay (v_P - v_S)   (masserror, V6:1096-1125);
  R_s(i,j)  = sum_k a_k (sc_P - sc_k) of coefsc + scalar (V6:795-812,
              861-964), a_w = h+(cx_W, Ds), a_e = h-(cx_P, Ds) (face fluxes
              unaveraged), amusc = amuf / prsc, plus the diffusion adds of
              every obstacle rectangle covering a neighbour (V6:939-963).

Penalised unknowns (u, v faces of obstacle cells, obstacle sc, obstacle p)
are fixed: the 1e30 penalty makes their derivatives O(1e-30) relative.
The unknowns x are the remaining interior u, v, p, sc of level 0; the
parameters theta are the inlet u profile (level-0 west face, rows 1..nfy),
the east boundary u, amuf and prsc (the coarse levels' copies of the
boundary values do not move the fixed point).  The objectives are the
device library's: J1 = mean p of the first interior column - mean p of the
last, J2 = sum over rows of areax * u_east * sc of the last cell.

Everything is differentiated by the complex step (x + i h, h = 1e-30, so
no cancellation: derivatives exact to rounding), with the branch of every
fmax / fabs frozen at the real value as the dual numbers of the device's
forward-mode library freeze it (fmax keeps its first argument at a tie).
The Jacobian K = dR/dx is assembled column group by column group (unknowns
of one field coloured by (i mod 3, j mod 3), no two of one colour in one
residual's stencil).  The adjoint system K^T lambda = dJ/dx is solved
directly (sparse LU); the pressure level of every connected fluid region
is a null vector of K and the region's summed continuity a null vector of
K^T, so one p unknown and one continuity row per region are dropped.  Then
dJ/dtheta = dJ/dtheta|_x - lambda^T dR/dtheta.  An inlet-row direction
also raises the east u by 1/nfy (a mass-consistent change, as the
device's gradient calls define it).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as spla

H = 1e-30   # complex step
PEN = 1e10  # a penalty array entry above this marks a fixed (penalised) unknown


def _abs(c):
    """fabs with the sign branch taken on the real part"""
    return np.where(c.real < 0.0, -c, c)


def _fmax(a, b):
    """fmax keeping its first argument at a tie (as the dual numbers do)"""
    return np.where(a.real >= b.real, a, b)


def hplus(c, d):
    return _fmax(_abs(c), d) + c


def hminus(c, d):
    return _fmax(_abs(c), d) - c


class SteadyAdjoint:
    """The level-0 steady residual of one problem at a converged state.

    state: dict of level-0 arrays u, v, p, sc (npy, npx) as the solver
    left them (boundary values included); pen: dict spu, spv, spsc, tag
    (level-0 penalty arrays of the oracle); obs_count: (npy, npx) count of
    obstacle rectangles covering each cell (the scalar adds are applied
    once per rectangle); dx, dy, ax, ay, amuf, prsc: the level-0 metrics
    and physics."""

    def __init__(self, state, pen, obs_count, dx, dy, ax, ay, amuf, prsc):
        self.npy, self.npx = state["u"].shape
        self.st = {k: np.asarray(state[k], dtype=np.float64) for k in ("u", "v", "p", "sc")}
        self.dx, self.dy, self.ax, self.ay = dx, dy, ax, ay
        self.amuf, self.prsc = amuf, prsc
        self.cnt = np.asarray(obs_count, dtype=np.float64)
        nx, ny = self.npx, self.npy
        J, I = np.mgrid[0:ny, 0:nx]
        self.J, self.I = J, I
        # active rows / unknowns
        self.mu = (I >= 1) & (I <= nx - 3) & (J >= 1) & (J <= ny - 2) & (pen["spu"] < PEN)
        self.mv = (I >= 1) & (I <= nx - 2) & (J >= 1) & (J <= ny - 3) & (pen["spv"] < PEN)
        cell = (I >= 1) & (I <= nx - 2) & (J >= 1) & (J <= ny - 2)
        self.mp = cell & (pen["tag"] > 0.5)
        self.ms = cell & (pen["spsc"] < PEN)
        self.masks = {"u": self.mu, "v": self.mv, "p": self.mp, "s": self.ms}
        self.nfy = ny - 2
        self._index()

    # ------------------------------------------------------------------
    def _index(self):
        off = 0
        self.col = {}
        for f in ("u", "v", "p", "s"):
            m = self.masks[f]
            idx = -np.ones(m.shape, dtype=np.int64)
            idx[m] = off + np.arange(int(m.sum()))
            off += int(m.sum())
            self.col[f] = idx
        self.n = off
        # rows: same layout (R_u on mu, R_v on mv, R_c on mp, R_s on ms)
        self.row = self.col

    def residual(self, u, v, p, s, inlet, uxp, amuf, prsc):
        """the four residual grids (complex), zero outside their rows"""
        nx, ny = self.npx, self.npy
        ax, ay, dx, dy = self.ax, self.ay, self.dx, self.dy
        u = u.astype(complex).copy()
        v = v.astype(complex)
        u[1:ny - 1, 0] = inlet
        u[1:ny - 1, nx - 2] = uxp
        cx = 0.5 * u * ax
        cy = 0.5 * v * ay
        Dx = amuf * (ax / dx)
        Dy = amuf * (ay / dy)
        R = {k: np.zeros((ny, nx), complex) for k in ("u", "v", "p", "s")}
        # u momentum
        j, i = slice(1, ny - 1), slice(1, nx - 2)
        sh = lambda a, dj, di: a[1 + dj:ny - 1 + dj, 1 + di:nx - 2 + di]
        cxm = 0.5 * (sh(cx, 0, 0) + sh(cx, 0, -1))
        cxp = 0.5 * (sh(cx, 0, 0) + sh(cx, 0, 1))
        cym = 0.5 * (sh(cy, -1, 0) + sh(cy, -1, 1))
        cyp = 0.5 * (sh(cy, 0, 0) + sh(cy, 0, 1))
        aw, ae = hplus(cxm, Dx), hminus(cxp, Dx)
        as_, an = hplus(cym, Dy), hminus(cyp, Dy)
        wall = amuf * (ay / dy)
        as_ = as_.copy()
        an = an.copy()
        as_[0, :] = as_[0, :] + wall
        an[-1, :] = an[-1, :] + wall
        uP = sh(u, 0, 0)
        ru = (aw * (uP - sh(u, 0, -1)) + ae * (uP - sh(u, 0, 1)) + as_ * (uP - sh(u, -1, 0))
              + an * (uP - sh(u, 1, 0)) - ax * (sh(p, 0, 0) - sh(p, 0, 1)))
        R["u"][j, i] = ru
        # v momentum
        j, i = slice(1, ny - 2), slice(1, nx - 1)
        sv = lambda a, dj, di: a[1 + dj:ny - 2 + dj, 1 + di:nx - 1 + di]
        cxm = 0.5 * (sv(cx, 0, -1) + sv(cx, 1, -1))
        cxp = 0.5 * (sv(cx, 0, 0) + sv(cx, 1, 0))
        cym = 0.5 * (sv(cy, 0, 0) + sv(cy, -1, 0))
        cyp = 0.5 * (sv(cy, 0, 0) + sv(cy, 1, 0))
        aw, ae = hplus(cxm, Dx).copy(), hminus(cxp, Dx).copy()
        as_, an = hplus(cym, Dy), hminus(cyp, Dy)
        side = amuf * (ay / dy)
        aw[:, 0] = aw[:, 0] + side
        ae[:, -1] = ae[:, -1] + side
        vP = sv(v, 0, 0)
        rv = (aw * (vP - sv(v, 0, -1)) + ae * (vP - sv(v, 0, 1)) + as_ * (vP - sv(v, -1, 0))
              + an * (vP - sv(v, 1, 0)) - ay * (sv(p, 0, 0) - sv(p, 1, 0)))
        R["v"][j, i] = rv
        # continuity
        j, i = slice(1, ny - 1), slice(1, nx - 1)
        sc_ = lambda a, dj, di: a[1 + dj:ny - 1 + dj, 1 + di:nx - 1 + di]
        R["p"][j, i] = ax * (sc_(u, 0, 0) - sc_(u, 0, -1)) + ay * (sc_(v, 0, 0) - sc_(v, -1, 0))
        # scalar
        amusc = amuf / prsc
        Dsx = amusc * (ax / dx)
        Dsy = amusc * (ay / dy)
        aw = hplus(sc_(cx, 0, -1), Dsx) + sc_(self.cnt, 0, -1) * Dsx
        ae = hminus(sc_(cx, 0, 0), Dsx) + sc_(self.cnt, 0, 1) * Dsx
        as_ = hplus(sc_(cy, -1, 0), Dsy) + sc_(self.cnt, -1, 0) * Dsy
        an = hminus(sc_(cy, 0, 0), Dsy) + sc_(self.cnt, 1, 0) * Dsy
        sP = sc_(s, 0, 0)
        R["s"][j, i] = (aw * (sP - sc_(s, 0, -1)) + ae * (sP - sc_(s, 0, 1))
                        + as_ * (sP - sc_(s, -1, 0)) + an * (sP - sc_(s, 1, 0)))
        for f in R:
            R[f][~self.masks[f]] = 0.0
        return R

    def _base(self):
        st = self.st
        return (st["u"].astype(complex), st["v"].astype(complex), st["p"].astype(complex),
                st["sc"].astype(complex), st["u"][1:-1, 0].astype(complex),
                complex(st["u"][1, self.npx - 2]), complex(self.amuf), complex(self.prsc))

    def residual_at_state(self):
        return {k: r.real for k, r in self.residual(*self._base()).items()}

    def _rowvec(self, R):
        out = np.zeros(self.n, complex)
        for f in ("u", "v", "p", "s"):
            m = self.masks[f]
            out[self.row[f][m]] = R[f][m]
        return out

    # stencil offsets (di, dj) of each unknown field in each residual
    _OFF = {
        ("u", "u"): [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)],
        ("u", "v"): [(0, -1), (1, -1), (0, 0), (1, 0)],
        ("u", "p"): [(0, 0), (1, 0)],
        ("v", "u"): [(-1, 0), (-1, 1), (0, 0), (0, 1)],
        ("v", "v"): [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)],
        ("v", "p"): [(0, 0), (0, 1)],
        ("p", "u"): [(0, 0), (-1, 0)],
        ("p", "v"): [(0, 0), (0, -1)],
        ("s", "u"): [(-1, 0), (0, 0)],
        ("s", "v"): [(0, -1), (0, 0)],
        ("s", "s"): [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)],
    }
    _ARG = {"u": 0, "v": 1, "p": 2, "s": 3}

    def jacobian(self):
        """K = dR/dx (sparse, n x n), by the complex step over coloured
        groups of unknowns"""
        rows, cols, vals = [], [], []
        base = list(self._base())
        for f in ("u", "v", "p", "s"):
            m = self.masks[f]
            for ci in range(3):
                for cj in range(3):
                    sel = m & (self.I % 3 == ci) & (self.J % 3 == cj)
                    if not sel.any():
                        continue
                    args = list(base)
                    a = args[self._ARG[f]].copy()
                    a[sel] += 1j * H
                    args[self._ARG[f]] = a
                    R = self.residual(*args)
                    for t in ("u", "v", "p", "s"):
                        offs = self._OFF.get((t, f))
                        if not offs:
                            continue
                        d = R[t].imag / H
                        jj, ii = np.nonzero(d != 0.0)
                        if len(jj) == 0:
                            continue
                        for di, dj in offs:
                            hit = ((ii + di) % 3 == ci) & ((jj + dj) % 3 == cj)
                            ju, iu = jj[hit] + dj, ii[hit] + di
                            ok = (ju >= 0) & (ju < self.npy) & (iu >= 0) & (iu < self.npx)
                            ju, iu, jr, ir = ju[ok], iu[ok], jj[hit][ok], ii[hit][ok]
                            c = self.col[f][ju, iu]
                            good = (c >= 0) & sel[ju, iu]
                            rows.append(self.row[t][jr[good], ir[good]])
                            cols.append(c[good])
                            vals.append(d[jr[good], ir[good]])
        K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(self.n, self.n)).tocsr()
        return K

    def param_jacobian(self):
        """dR/d(inlet rows) (n x nfy), dR/d(east u), dR/d(amuf), dR/d(prsc)"""
        base = list(self._base())
        nfy = self.nfy
        G = np.zeros((self.n, nfy))
        for c in range(3):
            args = list(base)
            inl = args[4].copy()
            rsel = (np.arange(nfy) % 3) == c
            inl[rsel] += 1j * H
            args[4] = inl
            d = self._rowvec(self.residual(*args)).imag / H
            # a residual row touched by an inlet row of this colour: rows of
            # the first column (u, p, s: row j; v: rows j and j + 1)
            for t in ("u", "v", "p", "s"):
                m = self.masks[t]
                jj, ii = np.nonzero(m & (self.I == 1))
                r = self.row[t][jj, ii]
                cand = [jj] if t != "v" else [jj, jj + 1]
                for jc in cand:
                    ok = (jc >= 1) & (jc <= nfy) & (((jc - 1) % 3) == c)
                    G[r[ok], jc[ok] - 1] += d[r[ok]]
        out = {"inlet": G}
        for k, pos in (("uxp", 5), ("amuf", 6), ("prsc", 7)):
            args = list(base)
            args[pos] = args[pos] + 1j * H
            out[k] = self._rowvec(self.residual(*args)).imag / H
        return out

    # ------------------------------------------------------------------
    def objective_grad(self, objective: int):
        """dJ/dx (length n) and the explicit dJ/d(east u)"""
        nx, ny = self.npx, self.npy
        g = np.zeros(self.n)
        djdu = 0.0
        if objective == 0:
            for j in range(1, ny - 1):
                for i, sgn in ((1, 1.0), (nx - 2, -1.0)):
                    c = self.col["p"][j, i]
                    if c >= 0:
                        g[c] += sgn / self.nfy
        else:
            uxp = self.st["u"][1, nx - 2]
            for j in range(1, ny - 1):
                c = self.col["s"][j, nx - 2]
                if c >= 0:
                    g[c] += self.ax * uxp
                djdu += self.ax * self.st["sc"][j, nx - 2]
        return g, djdu

    def objectives(self):
        nx, ny = self.npx, self.npy
        p, u, s = self.st["p"], self.st["u"], self.st["sc"]
        j1 = p[1:ny - 1, 1].sum() / self.nfy - p[1:ny - 1, nx - 2].sum() / self.nfy
        j2 = sum((self.ax * u[j, nx - 2]) * s[j, nx - 2] for j in range(1, ny - 1))
        return j1, j2

    def _reduce(self, K):
        """drop one p unknown and one continuity row per connected fluid
        region (the null spaces of K and K^T)"""
        cells = self.mp
        idx = -np.ones(cells.shape, np.int64)
        idx[cells] = np.arange(int(cells.sum()))
        a, b = [], []
        for f, di, dj in (("u", 1, 0), ("v", 0, 1)):
            m = self.masks[f]
            jj, ii = np.nonzero(m)
            c0, c1 = idx[jj, ii], idx[jj + dj, ii + di]
            ok = (c0 >= 0) & (c1 >= 0)
            a.append(c0[ok])
            b.append(c1[ok])
        nc = int(cells.sum())
        adj = sp.coo_matrix((np.ones(sum(len(x) for x in a)), (np.concatenate(a), np.concatenate(b))),
                            shape=(nc, nc))
        ncomp, lab = csg.connected_components(adj, directed=False)
        jj, ii = np.nonzero(cells)
        drop = set()
        for k in range(ncomp):
            first = np.nonzero(lab == k)[0][0]
            drop.add(int(self.col["p"][jj[first], ii[first]]))
        keep = np.array([q for q in range(self.n) if q not in drop])
        return keep, ncomp

    def solve_adjoint(self, objective: int, K=None):
        """lambda (length n, zero on the dropped rows) and the gradient:
        (J, dJ/d(inlet row r) with the east u raised by 1/nfy, dJ/damuf,
        dJ/dprsc)"""
        if K is None:
            K = self.jacobian()
        keep, _ = self._reduce(K)
        Kr = K[keep][:, keep]
        g, djdu = self.objective_grad(objective)
        lam = np.zeros(self.n)
        lam[keep] = spla.splu(Kr.T.tocsc()).solve(g[keep])
        P = self.param_jacobian()
        d_in = -(lam @ P["inlet"]) + (djdu - lam @ P["uxp"]) / self.nfy
        d_mu = -(lam @ P["amuf"])
        d_pr = -(lam @ P["prsc"])
        return lam, self.objectives()[objective], d_in, d_mu, d_pr

    def tangent(self, dtheta, K=None):
        """forward derivative of x along a parameter direction
        {inlet: nfy, uxp, amuf, prsc} (for cross-checks)"""
        if K is None:
            K = self.jacobian()
        keep, _ = self._reduce(K)
        P = self.param_jacobian()
        rhs = np.zeros(self.n)
        if dtheta.get("inlet") is not None:
            rhs += P["inlet"] @ np.asarray(dtheta["inlet"], float)
        for k in ("uxp", "amuf", "prsc"):
            rhs += P[k] * dtheta.get(k, 0.0)
        dx = np.zeros(self.n)
        dx[keep] = spla.splu(K[keep][:, keep].tocsc()).solve(-rhs[keep])
        return dx


def obstacle_count(prob, npx: int, npy: int) -> np.ndarray:
    """rectangles covering each level-0 cell, with the reference's index
    scaling (AllocateMemorySoA, V6:468-479): fac = 2^(ig_obs-1),
    is = (if + ibeg - 1) fac + 1, ie = (il + ibeg) fac, js = (jf - 1) fac + 1,
    je = jl fac"""
    cnt = np.zeros((npy, npx))
    if prob.ig_obs < 1:
        return cnt
    fac = 1 << (prob.ig_obs - 1)
    for a, b, c, d in prob.obstacles:
        is_, ie = (a + prob.ibeg_obs - 1) * fac + 1, (b + prob.ibeg_obs) * fac
        js, je = (c - 1) * fac + 1, d * fac
        cnt[js:je + 1, is_:ie + 1] += 1.0
    return cnt


def from_oracle(o, prob) -> SteadyAdjoint:
    """the steady residual of an oracle (orc.Oracle) at its current level-0
    state"""
    npx, npy = o.dims(0)
    st = {k: o.get(0, k) for k in ("u", "v", "p", "sc")}
    pen = {k: o.get(0, k) for k in ("spu", "spv", "spsc", "tag")}
    dx = prob.xl / (npx - 2)
    dy = prob.yl / (npy - 2)
    return SteadyAdjoint(st, pen, obstacle_count(prob, npx, npy), dx, dy, dy, dx, prob.amuf,
                         prob.prsc)


def from_state(state, pen, prob) -> SteadyAdjoint:
    """the same from given level-0 arrays (a device solver's state) and the
    oracle's penalty arrays of the same problem"""
    npy, npx = state["u"].shape
    dx = prob.xl / (npx - 2)
    dy = prob.yl / (npy - 2)
    return SteadyAdjoint(state, pen, obstacle_count(prob, npx, npy), dx, dy, dy, dx, prob.amuf,
                         prob.prsc)
