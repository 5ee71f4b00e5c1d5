"""CPU restatement of the learned correction's face features and training
targets (test infrastructure only: imported by tests/, never by the product
path).

The device builds, for an interior face of the coarse grid, 31 features in
FP32 and rounds them to BF16; this module repeats the same FP32 operations
in the same order with numpy (no fused multiply-adds, as the device code is
compiled with -ffp-contract=off), so the rows must agree bit for bit:

  stencil: c at the upwind cell + (s_n da, s_t dt, dk), da, dt, dk in -1..1,
           index (dk+1) 9 + (dt+1) 3 + (da+1), clamped to the interior,
           divided by S = max |c| over the 27 (rows with S <= 1e-30 are empty);
  27: |u_n| dt / dxp      (dxp = 0.5 / hdxp, hdxp = fp32(0.5 / dX_p(f-1)))
  28: |u_t| dt / dx_t     (u_t: mean of the 4 tangential faces around the face)
  29: w dt / dz           (mean of the 4 vertical faces of the two cells)
  30: v_s dt / dz         (bottom face of the upwind cell, this bin)
  31: 0
with s_n = sign(u_n), s_t = sign(u_t) (>= 0 counts as +), winds interpolated
in time as (1 - s) w1 + s w2. The target of a row is
  (0.5 (Phi_f(2f-1, 2t-1) + Phi_f(2f-1, 2t)) - Phi_c(f, t)) / (0.01 sgn S dxp)
from the fine (two steps) and coarse (one step) face mass fluxes.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def bf16_bits(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


class FaceFeatures:
    """Features of one coarse problem state (reference layouts, see
    fall3d_amd.problem.Problem); c is the device state (FP32 values)."""

    def __init__(self, prob, c, dt, stime=0.0):
        self.p = prob
        self.c = np.asarray(c, dtype=np.float32)
        self.dt = f32(dt)
        self.s = f32(stime)
        self.s1 = f32(1.0 - stime)
        g = prob.grid
        self.hdxp = [np.asarray(0.5 / np.asarray(g[k], dtype=np.float64), dtype=np.float32)
                     for k in ("dX1_p", "dX2_p", "dX3_p")]
        self.idxb = [np.asarray(1.0 / np.asarray(g[k], dtype=np.float64), dtype=np.float32)
                     for k in ("dX1_b", "dX2_b", "dX3_b")]
        P = prob
        self.u1, self.u2 = np.asarray(P.u1, np.float32), np.asarray(P.u2, np.float32)
        self.v1, self.v2 = np.asarray(P.v1, np.float32), np.asarray(P.v2, np.float32)
        self.w1, self.w2 = np.asarray(P.w1, np.float32), np.asarray(P.w2, np.float32)
        self.vs = np.asarray(P.vs, np.float32)

    def _lerp(self, a1, a2, idx):
        return f32(self.s1 * a1[idx]) + f32(self.s * a2[idx])

    def face(self, axis, b, f, tr, k):
        """(bf16 feature bits (32,), S, den = sgn S dxp) of one face; f is the
        face index (between cells f-1 and f), tr the cross index, k the level."""
        P = self.p
        nx, ny, nz = P.nx, P.ny, P.nz
        if axis == 0:
            iL, jL, iR, jR = f - 1, tr, f, tr
            un = self._lerp(self.u1, self.u2, (f, tr - 1, k - 1))
            ut = f32(0.0)
            for di in range(2):
                for dj in range(2):
                    ut = f32(ut + self._lerp(self.v1, self.v2, (tr + dj, f - 1 + di - 1, k - 1)))
        else:
            iL, jL, iR, jR = tr, f - 1, tr, f
            un = self._lerp(self.v1, self.v2, (f, tr - 1, k - 1))
            ut = f32(0.0)
            for dj in range(2):
                for di in range(2):
                    ut = f32(ut + self._lerp(self.u1, self.u2, (tr + di, f - 1 + dj - 1, k - 1)))
        ut = f32(ut * f32(0.25))
        wv = f32(0.0)
        for (i, j) in ((iL, jL), (iR, jR)):
            for dk in range(2):
                wv = f32(wv + self._lerp(self.w1, self.w2, (k + dk, i - 1, j - 1)))
        wv = f32(wv * f32(0.25))
        sgn = 1 if un >= 0 else -1
        tsgn = 1 if ut >= 0 else -1
        iU, jU = (iL, jL) if un >= 0 else (iR, jR)
        ta = 1 - axis
        tp = jU if axis == 0 else iU
        dxp = f32(f32(0.5) / self.hdxp[axis][f])
        cn = f32(f32(abs(un) * self.dt) / dxp)
        ct = f32(f32(abs(ut) * self.dt) * self.idxb[ta][tp])
        cw = f32(f32(wv * self.dt) * self.idxb[2][k])
        cs = f32(f32(self.vs[iU - 1, jU - 1, k, b] * self.dt) * self.idxb[2][k])
        vals = np.zeros(27, dtype=np.float32)
        for e in range(27):
            da, dt, dk = e % 3 - 1, (e // 3) % 3 - 1, e // 9 - 1
            sa, st = sgn * da, tsgn * dt
            i, j = (iU + sa, jU + st) if axis == 0 else (iU + st, jU + sa)
            i = min(max(i, 1), nx)
            j = min(max(j, 1), ny)
            kk = min(max(k + dk, 1), nz)
            vals[e] = self.c[i + 1, j + 1, kk + 1, b]
        S = f32(np.abs(vals).max())
        if not S > f32(1e-30):
            S = f32(0.0)
        feat = np.zeros(32, dtype=np.float32)
        feat[:27] = (vals / S).astype(np.float32) if S > 0 else 0.0
        feat[27:31] = (cn, ct, cw, cs)
        return bf16_bits(feat), S, f32(f32(f32(sgn) * S) * dxp)


def targets(ref, den, flux_fine, flux_coarse):
    """Training targets of collected rows from the face fluxes (float32).
    flux_*: (axis 0 array (nx+3, ny, nz, nb), axis 1 array (nx, ny+3, nz, nb))."""
    out = np.zeros(len(ref), dtype=np.float32)
    fx, fy = flux_fine
    cx, cy = flux_coarse
    for r, (q, dn) in enumerate(zip(ref, den)):
        axis, b, f, tr, k = q[0] & 1, q[0] >> 1, q[1], q[2], q[3]
        F = 2 * f - 1
        if axis == 0:
            pc = cx[f, tr - 1, k - 1, b]
            pf = f32(f32(0.0) + fx[F, 2 * tr - 2, k - 1, b])
            pf = f32(pf + fx[F, 2 * tr - 1, k - 1, b])
        else:
            pc = cy[tr - 1, f, k - 1, b]
            pf = f32(f32(0.0) + fy[2 * tr - 2, F, k - 1, b])
            pf = f32(pf + fy[2 * tr - 1, F, k - 1, b])
        out[r] = f32(f32(f32(0.5) * pf) - pc) / f32(f32(0.01) * dn) if dn != 0 else 0.0
    return out
