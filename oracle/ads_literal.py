"""Literal pure-Python restatement of the reference ADS sweeps (tiny inputs only).

TEST INFRASTRUCTURE (see ads_oracle.c). This second restatement follows the
reference's own control structure rather than the C oracle's per-line one:
each sweep runs stage 1 over ALL lines of a bin into full-size c0/new_c
work volumes, then stage 2 over all lines, and so on
(Sources/mod_ADS.F90:313-399, 502-590, 710-798), with Fortran index bounds
kept through small offset-array helpers. Agreement of the two
restatements on random tiny problems pins the index bookkeeping of the C
oracle (a transposed layout, an off-by-one halo or a misplaced ghost copy
fails it); the closed-form tests pin the arithmetic itself.

Pure Python loops: keep inputs to a few hundred cells.
"""
from __future__ import annotations

EPSILON = 1e-12  # Sources/mod_KindType.F90:42


class FArr:
    """Fortran-style array view over a flat numpy array with explicit bounds."""

    def __init__(self, data, bounds):
        self.d = data.reshape(-1, order="F")
        self.lo = [b[0] for b in bounds]
        self.ext = [b[1] - b[0] + 1 for b in bounds]

    def _off(self, idx):
        o, s = 0, 1
        for i, lo, e in zip(idx, self.lo, self.ext):
            assert lo <= i < lo + e, (idx, self.lo, self.ext)
            o += (i - lo) * s
            s *= e
        return o

    def __getitem__(self, idx):
        return float(self.d[self._off(idx)])

    def __setitem__(self, idx, v):
        self.d[self._off(idx)] = v


def _limiter(kind, c, ips, ipe):
    """r_minmod / r_sbee / r_ospre, Sources/mod_ADS.F90:1124-1223."""
    r = {}
    for i in range(ips - 1, ipe + 2):
        if abs(c[i + 1] - c[i]) < EPSILON:
            if kind == 1:
                r[i] = 1.0
            elif abs(c[i] - c[i - 1]) < EPSILON:
                r[i] = 0.0
            else:
                r[i] = 2.0 if kind == 2 else 1.5
        else:
            rr = (c[i] - c[i - 1]) / (c[i + 1] - c[i])
            if kind == 1:
                r[i] = max(0.0, min(1.0, rr))
            elif kind == 2:
                r[i] = max(0.0, min(2.0 * rr, 1.0), min(rr, 2.0))
            else:
                r[i] = 1.5 * (rr * rr + rr) / (rr * rr + rr + 1.0)
    r[ips - 2] = r[ips - 1]
    r[ipe + 2] = r[ipe + 1]
    return r


def kt_rhs(kind, c, dxp, dxb, u, k, ips, ipe):
    """KT_RHS, Sources/mod_ADS.F90:1057-1120. c, dxp, dxb, u, k are dicts keyed by
    the Fortran index. Returns (rhs dict over ips..ipe, flux tuple)."""
    r = _limiter(kind, c, ips, ipe)
    cr, cl = {}, {}
    for i in range(ips - 1, ipe + 2):
        cr[i] = c[i] - 0.5 * r[i] * (c[i + 1] - c[i])
        cl[i] = c[i - 1] + 0.5 * r[i - 1] * (c[i] - c[i - 1])
    rhs = {}
    flux = [0.0, 0.0]
    for i in range(ips, ipe + 1):
        F_ml = cl[i] * u[i]
        F_mr = cr[i] * u[i]
        F_pl = cl[i + 1] * u[i + 1]
        F_pr = cr[i + 1] * u[i + 1]
        F_m = 0.5 * (F_ml + F_mr) - 0.5 * abs(u[i]) * (cr[i] - cl[i])
        F_p = 0.5 * (F_pl + F_pr) - 0.5 * abs(u[i + 1]) * (cr[i + 1] - cl[i + 1])
        P_m = 0.5 * (k[i] + k[i - 1]) * (c[i] - c[i - 1]) / dxp[i - 1]
        P_p = 0.5 * (k[i + 1] + k[i]) * (c[i + 1] - c[i]) / dxp[i]
        rhs[i] = -1.0 * (F_p - F_m) / dxb[i] + (P_p - P_m) / dxb[i]
        if i == ips:
            flux[0] = F_m + P_m
        if i == ipe:
            flux[1] = F_p + P_p
    return rhs, flux


def _solve(dirn, prob, kind, tm, dt, c_np, kd_np, vel, flux, acum_np=None, w2=None, stime=0.0):
    """Shared body of ADS_solve_along_{x,y,z}. dirn 0/1/2. `vel` is u, v or w1
    (w1/w2 interpolated with stime minus vs for z). flux: 2-list accumulated."""
    nx, ny, nz, nb = prob.nx, prob.ny, prob.nz, prob.nb
    c = FArr(c_np, [(-1, nx + 2), (-1, ny + 2), (-1, nz + 2), (1, nb)])
    g = prob.grid
    n = (nx, ny, nz)[dirn]
    t1n, t2n = [(ny, nz), (nx, nz), (nx, ny)][dirn]
    dxp = {i: float(g[("dX1_p", "dX2_p", "dX3_p")[dirn]][i + 1]) for i in range(-1, n + 3)}
    dxb = {i: float(g[("dX1_b", "dX2_b", "dX3_b")[dirn]][i]) for i in range(0, n + 3)}
    gb = [{i: float(g[name][i]) for i in range(0, m + 3)} for name, m in
          (("dX1_b", nx), ("dX2_b", ny), ("dX3_b", nz))]
    kd = FArr(kd_np, [(-1, n + 2), (1, t1n), (1, t2n)])
    V = FArr(vel, [(0, n + 2), (1, t1n), (1, t2n)])
    if dirn == 2:
        V2 = FArr(w2, [(0, n + 2), (1, t1n), (1, t2n)])
        VS = FArr(prob.vs, [(1, nx), (1, ny), (0, nz + 2), (1, nb)])
        acum = FArr(acum_np, [(-1, nx + 2), (-1, ny + 2), (1, nb)])

    def cell(t1, t2, p):  # (i,j,k) of point p on line (t1,t2)
        return [(p, t1, t2), (t1, p, t2), (t1, t2, p)][dirn]

    def area(t1, t2):
        return [gb[2][t2] * gb[1][t1], gb[2][t2] * gb[0][t1], gb[1][t2] * gb[0][t1]][dirn]

    for ibin in range(1, nb + 1):
        # velocity along each line (for z: w_eff per bin, mod_ADS.F90:667-672)
        vline = {}
        for t2 in range(1, t2n + 1):
            for t1 in range(1, t1n + 1):
                if dirn == 2:
                    vline[t1, t2] = {f: ((1.0 - stime) * V[f, t1, t2] + stime * V2[f, t1, t2]) - VS[t1, t2, f, ibin]
                                     for f in range(0, n + 3)}
                else:
                    vline[t1, t2] = {f: V[f, t1, t2] for f in range(0, n + 3)}
        # freeflow, Sources/mod_ADS.F90:1227-1326 (physical boundaries on both sides)
        for t2 in range(1, t2n + 1):
            for t1 in range(1, t1n + 1):
                vl = vline[t1, t2]
                first = c[(*cell(t1, t2, 1), ibin)]
                g0 = (2.0 * 0.0 - first) if vl[1] > 0.0 else first
                c[(*cell(t1, t2, 0), ibin)] = g0
                c[(*cell(t1, t2, -1), ibin)] = g0
                last = c[(*cell(t1, t2, n), ibin)]
                gn = (2.0 * 0.0 - last) if vl[n + 1] < 0.0 else last
                c[(*cell(t1, t2, n + 1), ibin)] = gn
                c[(*cell(t1, t2, n + 2), ibin)] = gn
        lines = [(t1, t2) for t2 in range(1, t2n + 1) for t1 in range(1, t1n + 1)]
        ini = {L: {p: c[(*cell(*L, p), ibin)] for p in range(-1, n + 3)} for L in lines}
        kdl = {L: {p: kd[p, L[0], L[1]] for p in range(-1, n + 3)} for L in lines}
        if tm == 1:
            for L in lines:
                ki, fl = kt_rhs(kind, ini[L], dxp, dxb, vline[L], kdl[L], 1, n)
                for p in range(1, n + 1):
                    c[(*cell(*L, p), ibin)] = ini[L][p] + dt * ki[p]
                _acc(dirn, flux, fl, area(*L), dt, acum if dirn == 2 else None, L, ibin)
            continue
        c0 = {L: dict(ini[L]) for L in lines}
        new = {L: dict(ini[L]) for L in lines}
        for L in lines:  # stage 1
            ki, _ = kt_rhs(kind, ini[L], dxp, dxb, vline[L], kdl[L], 1, n)
            for p in range(1, n + 1):
                c0[L][p] = c0[L][p] + 0.5 * dt * ki[p]
                new[L][p] = new[L][p] + dt * ki[p] / 6.0
        for fac in (0.5, 1.0):  # stages 2 and 3
            stage = {}
            for L in lines:
                ki, _ = kt_rhs(kind, c0[L], dxp, dxb, vline[L], kdl[L], 1, n)
                stage[L] = ki
            for L in lines:
                for p in range(1, n + 1):
                    c0[L][p] = ini[L][p] + fac * dt * stage[L][p] if fac != 1.0 else ini[L][p] + dt * stage[L][p]
                    new[L][p] = new[L][p] + dt * stage[L][p] / 3.0
        for L in lines:  # stage 4
            ki, fl = kt_rhs(kind, c0[L], dxp, dxb, vline[L], kdl[L], 1, n)
            for p in range(1, n + 1):
                c[(*cell(*L, p), ibin)] = new[L][p] + dt * ki[p] / 6.0
            _acc(dirn, flux, fl, area(*L), dt, acum if dirn == 2 else None, L, ibin)


def _acc(dirn, flux, fl, area, dt, acum, L, ibin):
    flux[0] = flux[0] + fl[0] * area * dt
    flux[1] = flux[1] + fl[1] * area * dt
    if acum is not None:
        t1, t2 = L
        acum[t1, t2, ibin] = acum[t1, t2, ibin] + abs(fl[0]) * dt


def solve_x(prob, kind, tm, dt, c, k1, u, flux):
    _solve(0, prob, kind, tm, dt, c, k1, u, flux)


def solve_y(prob, kind, tm, dt, c, k2, v, flux):
    _solve(1, prob, kind, tm, dt, c, k2, v, flux)


def solve_z(prob, kind, tm, dt, c, acum, k3, w1, w2, stime, flux):
    _solve(2, prob, kind, tm, dt, c, k3, w1, flux, acum_np=acum, w2=w2, stime=stime)
