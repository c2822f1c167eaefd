"""FP64 CPU reference gradient of one step with frozen neighbor lists and labels.

TEST INFRASTRUCTURE: only tests/ may import this module.

The reference has no adjoint; this module states the reverse mode of the
reference's step in plain numpy over CSR pair arrays (scatter form, np.add.at),
with its forward intermediates taken from the FP64 oracle stages
(oracle/lsmps_oracle.cpp) and the PPE solved directly (scipy SuperLU) so that the
frozen step is a smooth function for central finite differences.  The stage
formulas differentiated are those of
  src/moment_matrix/moment_matrix.cu:211-309   (moments + regularisation + inverse)
  src/provision/explicit_update.cu:148-270     (velocity Laplacian, u*)
  src/ppe/ppe_matrix.cu:203-361                (PPE rows, divergence, wall term)
  src/correction/pressure_correction.cu:188-515 (clamp, gradient, shifting,
                                                update, anti-penetration, smoothing)
Parity of this module is pinned by central finite differences of
`FrozenStep.forward` (tests/test_adjoint_oracle.py); the GPU adjoint
(lsmps_adjoint_step) is compared with it.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

INNER, NEAR_SURFACE, SURFACE, SPLASH = 0, 1, 2, 3


def _weight(r, re, quad):
    t = 1.0 - r / re
    return np.where(r < re, t * t if quad else t, 0.0)


def _slope(r, re, quad):
    t = 1.0 - r / re
    return np.where(r < re, -2.0 * t / re if quad else -1.0 / re + 0.0 * r, 0.0)


def _basis(d, re):
    s = d / re
    return np.stack([s[0], s[1], s[2], s[0] * s[0], s[1] * s[1], s[2] * s[2], s[0] * s[1],
                     s[1] * s[2], s[2] * s[0]])


def _basis_jt(d, re, u):
    """J_b(d)^T u for u of shape (9, m)."""
    s = d / re
    return np.stack([u[0] + 2 * s[0] * u[3] + s[1] * u[6] + s[2] * u[8],
                     u[1] + 2 * s[1] * u[4] + s[0] * u[6] + s[2] * u[7],
                     u[2] + 2 * s[2] * u[5] + s[1] * u[7] + s[0] * u[8]]) / re


def _neumann(d, nrm, re):
    return np.stack([nrm[0], nrm[1], nrm[2], 2 * d[0] * nrm[0] / re, 2 * d[1] * nrm[1] / re,
                     2 * d[2] * nrm[2] / re, (d[1] * nrm[0] + d[0] * nrm[1]) / re,
                     (d[2] * nrm[1] + d[1] * nrm[2]) / re, (d[0] * nrm[2] + d[2] * nrm[0]) / re])


def _neumann_jt(nrm, re, u):
    return np.stack([2 * nrm[0] * u[3] + nrm[1] * u[6] + nrm[2] * u[8],
                     2 * nrm[1] * u[4] + nrm[0] * u[6] + nrm[2] * u[7],
                     2 * nrm[2] * u[5] + nrm[1] * u[7] + nrm[0] * u[8]]) / re


def _scatter(n, idx, vals):
    """sum of vals (k, m) into (k, n) at columns idx."""
    vals = np.atleast_2d(vals)
    return np.stack([np.bincount(idx, weights=vals[k], minlength=n).astype(np.float64)
                     for k in range(vals.shape[0])])


def _op_rows(minv):
    """Operator rows 0, 1, 2 and 3+4+5 of each inverse: (n, 4, 9)."""
    return np.stack([minv[:, 0, :], minv[:, 1, :], minv[:, 2, :],
                     minv[:, 3, :] + minv[:, 4, :] + minv[:, 5, :]], axis=1)


class FrozenStep:
    """One step of `case` with the neighbor lists `nb` and labels `st` held fixed."""

    def __init__(self, orc, case, nb=None, st=None, max_neighbors=256):
        self.orc = orc
        self.abi = orc.abi
        self.case = case
        self.max_neighbors = max_neighbors
        if nb is None:
            nb = orc.neighbors(case, max_neighbors)
        if st is None:
            st = orc.classify(case, nb, diagnostics=False)["surface_type"]
        self.nb = nb
        self.st = np.ascontiguousarray(st, np.int32)
        n = case.n_fluid
        f_off = nb["fluid_offsets"].astype(np.int64)
        w_off = nb["wall_offsets"].astype(np.int64)
        self.I = np.repeat(np.arange(n), np.diff(f_off))
        self.J = nb["fluid_indices"][:f_off[-1]].astype(np.int64)
        self.IW = np.repeat(np.arange(n), np.diff(w_off))
        self.JW = nb["wall_indices"][:w_off[-1]].astype(np.int64)
        self.w_off = w_off

    # -- forward ------------------------------------------------------------
    def _case(self, x, v, g):
        c = self.case.copy()
        for k, key in enumerate(("x", "y", "z")):
            c.fluid[key] = np.ascontiguousarray(x[k], np.float64).copy()
        for k, key in enumerate(("vx", "vy", "vz")):
            c.fluid[key] = np.ascontiguousarray(v[k], np.float64).copy()
        c.fluid["surface_type"] = self.st.copy()
        for k in range(3):
            c.config.gravity[k] = float(g[k])
        return c

    def state(self):
        f = self.case.fluid
        return (np.stack([f["x"], f["y"], f["z"]]), np.stack([f["vx"], f["vy"], f["vz"]]),
                np.array(list(self.case.config.gravity)))

    def forward(self, x, v, g, pressure=None):
        """Positions / velocities after the step; `pressure` replaces the PPE
        solution (before the clamp) when given."""
        a = self.abi
        o = self.orc
        c = self._case(x, v, g)
        n = c.n_fluid
        vel = o.moments(c, self.nb, a.MOMENT_VELOCITY_WALL_DIRICHLET_A)
        pre = o.moments(c, self.nb, a.MOMENT_PRESSURE_WALL_NEUMANN_A)
        tmp = o.provision(c, self.nb, vel["inverse"], surface_type=self.st)
        ppe = o.assemble(c, self.nb, tmp, vel["inverse"], pre["inverse"], surface_type=self.st,
                         neighbor_capacity=n * self.max_neighbors)
        A = sp.csr_matrix((ppe["values"], ppe["col_indices"], ppe["row_offsets"]), shape=(n, n))
        p = spla.spsolve(A.tocsc(), ppe["rhs"]) if pressure is None else np.array(pressure)
        cor = o.correction(c, self.nb, tmp, np.maximum(p, 0.0), pre["inverse"],
                           surface_type=self.st)
        f = cor["fluid"]
        return {"case": c, "vel": vel, "pre": pre, "tmp": tmp, "ppe": ppe, "A": A, "p": p,
                "cor": cor, "x": np.stack([c.fluid["x"], c.fluid["y"], c.fluid["z"]]),
                "v": np.stack([c.fluid["vx"], c.fluid["vy"], c.fluid["vz"]]),
                "x_new": np.stack([f["x"], f["y"], f["z"]]),
                "v_new": np.stack([f["vx"], f["vy"], f["vz"]])}

    # -- reverse ------------------------------------------------------------
    def adjoint(self, fw, vbar, xbar_out=None):
        """dL/dx, dL/dv (start of step), dL/dp (clamped pressure read by the
        correction) and dL/dg for upstream dL/dv_new = vbar (3, n) and
        dL/dx_new = xbar_out."""
        c = fw["case"]
        cfg = c.config
        n = c.n_fluid
        re = cfg.support_radius
        quad = cfg.weight_kernel == self.abi.WEIGHT_QUADRATIC
        dt, rho, nu = cfg.time_step, cfg.density, cfg.kinematic_viscosity
        wsc = cfg.lsmps_wall_weight_scale
        lam_reg = cfg.lsmps_regularization
        g = np.array(list(cfg.gravity))
        x, v = fw["x"], fw["v"]
        W = c.walls
        wp = np.stack([W["x"], W["y"], W["z"]])
        wn = np.stack([W["normal_x"], W["normal_y"], W["normal_z"]])
        wv = np.stack([W["vx"], W["vy"], W["vz"]])
        st = self.st
        splash = st == SPLASH
        dirichlet = (st == SURFACE) | splash
        tmp = fw["tmp"]
        tv = np.stack([tmp["vx"], tmp["vy"], tmp["vz"]])
        tw = np.stack([tmp["wall_vx"], tmp["wall_vy"], tmp["wall_vz"]])
        grad = np.stack(fw["cor"]["grad"])
        ps = np.stack(fw["cor"]["ps"])
        p_raw = fw["p"]
        pc = np.maximum(p_raw, 0.0)
        I, J, IW, JW = self.I, self.J, self.IW, self.JW
        xbar = np.zeros((3, n))
        vbar_in = np.zeros((3, n))
        gbar = np.zeros(3)

        # forward replay: update, anti-penetration
        nv = tv.copy()
        nv[:, ~splash] -= dt / rho * grad[:, ~splash]
        xp = x + (0.5 * dt * (v + nv) + ps)
        clearance = cfg.wall_clearance_ratio * cfg.particle_spacing
        counts = np.diff(self.w_off)
        kmax = int(counts.max()) if n else 0
        xpp, vpp = xp.copy(), nv.copy()
        pos_act, vel_act, widx = [], [], []
        if clearance > 0.0:
            for k in range(kmax):
                rows = np.nonzero(counts > k)[0]
                j = self.JW[self.w_off[rows] + k]
                nrm = wn[:, j]
                sd = ((xpp[:, rows] - wp[:, j]) * nrm).sum(0)
                act = sd < clearance
                ra, ja = rows[act], j[act]
                na = wn[:, ja]
                xpp[:, ra] += (clearance - sd[act]) * na
                vn = (vpp[:, ra] * na).sum(0)
                va = vn < 0.0
                vpp[:, ra[va]] -= vn[va] * na[:, va]
                pos_act.append(ra)
                vel_act.append(ra[va])
                widx.append((ja, ja[va]))

        # smoothing at the corrected positions
        strength = cfg.velocity_smoothing_strength
        alpha = min(max(strength, 0.0), 1.0)
        vb_pp = vbar.copy()
        xb_pp = np.zeros((3, n)) if xbar_out is None else np.array(xbar_out, dtype=np.float64)
        if strength > 0.0:
            d = xpp[:, J] - xpp[:, I]
            r = np.sqrt((d * d).sum(0))
            w = np.where(r > 0.0, _weight(r, re, quad), 0.0)
            use = (w > 0.0) & ~splash[J] & ~splash[I]
            Wsum = np.zeros(n)
            np.add.at(Wsum, I[use], w[use])
            msum = _scatter(n, I[use], w[use] * vpp[:, J[use]])
            elig = (Wsum > 0.0) & ~splash
            m = np.where(elig, msum / np.where(elig, Wsum, 1.0), 0.0)
            vb_pp[:, elig] *= (1.0 - alpha)
            u = use & elig[I]
            Iu, Ju = I[u], J[u]
            coef_v = alpha * w[u] / Wsum[Iu]
            vb_pp += _scatter(n, Ju, coef_v * vbar[:, Iu])
            coeff = alpha / Wsum[Iu] * ((vpp[:, Ju] - m[:, Iu]) * vbar[:, Iu]).sum(0)
            gpair = coeff * _slope(r[u], re, quad) / r[u] * d[:, u]
            xb_pp += _scatter(n, Ju, gpair) - _scatter(n, Iu, gpair)

        # anti-penetration in reverse
        xb_p, nvb = xb_pp.copy(), vb_pp.copy()
        for k in reversed(range(len(pos_act))):
            ra, rv = pos_act[k], vel_act[k]
            ja, jv = widx[k]
            na = wn[:, ja]
            xb_p[:, ra] -= (xb_p[:, ra] * na).sum(0) * na
            nvv = wn[:, jv]
            nvb[:, rv] -= (nvb[:, rv] * nvv).sum(0) * nvv

        # update
        nvb += 0.5 * dt * xb_p
        ubar = nvb.copy()
        gradbar = np.where(splash, 0.0, -dt / rho * nvb)
        psbar = xb_p.copy()
        xbar += xb_p
        vbar_in += 0.5 * dt * xb_p

        # particle shifting
        m_d = cfg.ps_min_distance_ratio * cfg.particle_spacing
        pmax = cfg.ps_max_displacement_ratio * cfg.particle_spacing
        scale = cfg.ps_displacement_scale
        if scale > 0.0:
            active = ~splash
            d = x[:, I] - x[:, J]
            r = np.sqrt((d * d).sum(0))
            u = (r > 0.0) & (r < m_d) & active[I]
            ru = np.where(u, r, 1.0)
            sh = _scatter(n, I[u], (scale * (m_d - r[u]) / r[u]) * d[:, u])
            na = _scatter(n, I[u], d[:, u] / r[u])
            dw_ = x[:, IW] - wp[:, JW]
            rw = np.sqrt((dw_ * dw_).sum(0))
            uw = (rw > 0.0) & (rw < m_d) & active[IW]
            sh += _scatter(n, IW[uw], scale * (m_d - rw[uw]) * wn[:, JW[uw]])
            nn = np.sqrt((na * na).sum(0))
            proj = (st == SURFACE) & (nn > 0.0)
            nh = np.where(proj, na / np.where(proj, nn, 1.0), 0.0)
            comp = (sh * nh).sum(0)
            y = np.where(proj, sh - comp * nh, sh)
            norm = np.sqrt((y * y).sum(0))
            clip = (norm > pmax) & (norm > 0.0) & active
            yb = np.where(active, psbar, 0.0)
            yu = (yb * y).sum(0) / np.where(clip, norm * norm, 1.0)
            yb = np.where(clip, pmax / np.where(clip, norm, 1.0) * (yb - yu * y), yb)
            ybn = (yb * nh).sum(0)
            shb = np.where(proj, yb - ybn * nh, yb)
            nhb = -ybn * sh - comp * yb
            t = (nhb * nh).sum(0)
            nab = np.where(proj, (nhb - t * nh) / np.where(proj, nn, 1.0), 0.0)
            du = d[:, u]
            rr = ru[u]
            sI = shb[:, I[u]]
            nI = nab[:, I[u]]
            ds = (du * sI).sum(0)
            dn = (du * nI).sum(0)
            gd = scale * ((m_d / rr - 1.0) * sI - m_d * ds * du / rr ** 3) + (nI - dn * du / rr ** 2) / rr
            xbar += _scatter(n, I[u], gd) - _scatter(n, J[u], gd)
            dwu = dw_[:, uw]
            gw = -scale * (wn[:, JW[uw]] * shb[:, IW[uw]]).sum(0) / rw[uw] * dwu
            xbar += _scatter(n, IW[uw], gw)

        # pressure gradient (clamped pressure), operator rows of the pressure kind
        pre_inv = fw["pre"]["inverse"].transpose(0, 2, 1)
        vel_inv = fw["vel"]["inverse"].transpose(0, 2, 1)
        pops = _op_rows(pre_inv)
        vops = _op_rows(vel_inv)
        opbar_p = np.zeros((n, 4, 9))
        opbar_v = np.zeros((n, 4, 9))
        ns = ~splash
        d = x[:, J] - x[:, I]
        r = np.sqrt((d * d).sum(0))
        wc = np.where(r > 0.0, _weight(r, re, quad), 0.0)
        dwc = np.where(r > 0.0, _slope(r, re, quad), 0.0)
        b = _basis(d, re)
        u = (wc > 0.0) & ns[I]
        rhs = _scatter(n, I[u], wc[u] * b[:, u] * (pc[J[u]] - pc[I[u]]))
        dwall = wp[:, JW] - x[:, IW]
        rwall = np.sqrt((dwall * dwall).sum(0))
        wcw = wsc * np.where(rwall > 0.0, _weight(rwall, re, quad), 0.0)
        dwcw = wsc * np.where(rwall > 0.0, _slope(rwall, re, quad), 0.0)
        nW = wn[:, JW]
        qW = _neumann(dwall, nW, re)
        ngW = (g[:, None] * nW).sum(0)
        uw = (wcw > 0.0) & ns[IW]
        rhs += _scatter(n, IW[uw], wcw[uw] * qW[:, uw] * (re * rho * ngW[uw]))
        gb = np.where(ns, gradbar, 0.0)
        for q in range(3):
            opbar_p[:, q, :] = (gb[q] / re)[:, None] * rhs.T
        rb = np.einsum("qn,nqk->kn", gb / re, pops[:, :3, :])
        dp = pc[J[u]] - pc[I[u]]
        rbI = rb[:, I[u]]
        ub_ = (b[:, u] * rbI).sum(0)
        gp = dp * (dwc[u] / r[u] * ub_ * d[:, u] + wc[u] * _basis_jt(d[:, u], re, rbI))
        xbar += _scatter(n, J[u], gp) - _scatter(n, I[u], gp)
        pbar = np.zeros(n)
        np.add.at(pbar, J[u], wc[u] * ub_)
        np.add.at(pbar, I[u], -wc[u] * ub_)
        rbW = rb[:, IW[uw]]
        uq = (qW[:, uw] * rbW).sum(0)
        wr = re * rho * ngW[uw]
        gxw = wr * (dwcw[uw] / rwall[uw] * uq * dwall[:, uw] + wcw[uw] * _neumann_jt(nW[:, uw], re, rbW))
        xbar -= _scatter(n, IW[uw], gxw)
        gbar += (wcw[uw] * uq * re * rho * nW[:, uw]).sum(1)

        # clamp + PPE solve: lambda = A^-T (dL/dp on rows with p > 0)
        psolve = np.where(p_raw > 0.0, pbar, 0.0)
        lam = spla.spsolve(fw["A"].T.tocsc(), psolve)

        # PPE assembly (non-Dirichlet rows)
        nd = ~dirichlet
        coef = 2.0 / (re * re * rho)
        w = _weight(r, re, quad)
        dw = _slope(r, re, quad)
        ua = (w > 0.0) & nd[I]
        rr = np.zeros((3, 9, n))
        for q in range(3):
            rr[q] = _scatter(n, I[ua], w[ua] * b[:, ua] * (tv[q, J[ua]] - tv[q, I[ua]]))
        bw_ = np.where(True, _weight(rwall, re, quad), 0.0)
        dbw = _slope(rwall, re, quad)
        bW = _basis(dwall, re)
        uwa = (bw_ > 0.0) & nd[IW]
        for q in range(3):
            rr[q] += _scatter(n, IW[uwa], wsc * bw_[uwa] * bW[:, uwa] *
                              (tw[q, JW[uwa]] - tv[q, IW[uwa]]))
        divb = np.where(nd, lam / dt, 0.0)
        wallb = np.where(nd, -lam, 0.0)
        for q in range(3):
            opbar_v[:, q, :] = (divb / re)[:, None] * rr[q].T
        rbar = np.einsum("n,nqk->qkn", divb / re, vops[:, :3, :])  # (3, 9, n)
        Lp = pops[:, 3, :].T  # (9, n)
        Lb = np.zeros((9, n))
        Ia, Ja = I[ua], J[ua]
        da, ra_, wa, dwa = d[:, ua], r[ua], w[ua], dw[ua]
        dwra = np.where(ra_ > 0.0, dwa / np.where(ra_ > 0.0, ra_, 1.0), 0.0)
        ba = b[:, ua]
        LI = Lp[:, Ia]
        lb = (LI * ba).sum(0)
        offb = -lam[Ia] * (p_raw[Ja] - p_raw[Ia]) * coef
        Lb += _scatter(n, Ia, offb * wa * ba)
        gpair = offb * (dwra * lb * da + wa * _basis_jt(da, re, LI))
        ubar_loc = np.zeros((3, n))
        for q in range(3):
            rq = rbar[q][:, Ia]
            bq = (ba * rq).sum(0)
            du = tv[q, Ja] - tv[q, Ia]
            gpair += du * (dwra * bq * da + wa * _basis_jt(da, re, rq))
            np.add.at(ubar_loc[q], Ja, wa * bq)
            np.add.at(ubar_loc[q], Ia, -wa * bq)
        xbar += _scatter(n, Ja, gpair) - _scatter(n, Ia, gpair)
        IWa, JWa = IW[uwa], JW[uwa]
        dW, rW, bwa = dwall[:, uwa], rwall[uwa], bw_[uwa]
        dbwr = np.where(rW > 0.0, dbw[uwa] / np.where(rW > 0.0, rW, 1.0), 0.0)
        bWa = bW[:, uwa]
        for q in range(3):
            rq = rbar[q][:, IWa]
            bq = (bWa * rq).sum(0)
            du = tw[q, JWa] - tv[q, IWa]
            xbar -= _scatter(n, IWa, du * wsc * (dbwr * bq * dW + bwa * _basis_jt(dW, re, rq)))
            np.add.at(ubar_loc[q], IWa, -wsc * bwa * bq)
            gbar[q] += (dt * wsc * bwa * bq).sum()
        nWa = wn[:, JWa]
        ngWa = (g[:, None] * nWa).sum(0)
        qWa = _neumann(dW, nWa, re)
        LIW = Lp[:, IWa]
        lq = (LIW * qWa).sum(0)
        cw = 2.0 * wsc / re
        wbI = wallb[IWa]
        gbar += (wbI * cw * bwa * lq * nWa).sum(1)
        Lb += _scatter(n, IWa, wbI * cw * bwa * ngWa * qWa)
        xbar -= _scatter(n, IWa, wbI * cw * ngWa * (dbwr * lq * dW + bwa * _neumann_jt(nWa, re, LIW)))
        opbar_p[:, 3, :] = Lb.T
        ubar += ubar_loc

        # provision: u* = v + dt (nu_eff lap + g)
        gbar += dt * ubar.sum(1)
        nu_eff = np.where(splash, 0.0, nu)
        sbar = 2.0 / (re * re) * dt * nu_eff * ubar  # (3, n)
        vbar_in += ubar
        Lv = vops[:, 3, :].T
        up = w > 0.0
        Ip, Jp = I[up], J[up]
        dp_, rp, wp_ = d[:, up], r[up], w[up]
        dwrp = np.where(rp > 0.0, dw[up] / np.where(rp > 0.0, rp, 1.0), 0.0)
        bp = b[:, up]
        LvI = Lv[:, Ip]
        lbp = (LvI * bp).sum(0)
        cc = ((v[:, Jp] - v[:, Ip]) * sbar[:, Ip]).sum(0)
        Lvb = _scatter(n, Ip, wp_ * cc * bp)
        gpair = cc * (dwrp * lbp * dp_ + wp_ * _basis_jt(dp_, re, LvI))
        xbar += _scatter(n, Jp, gpair) - _scatter(n, Ip, gpair)
        vbar_in += _scatter(n, Jp, wp_ * lbp * sbar[:, Ip]) - _scatter(n, Ip, wp_ * lbp * sbar[:, Ip])
        upw = wsc * bw_ > 0.0
        IWp, JWp = IW[upw], JW[upw]
        dWp, rWp, bwp = dwall[:, upw], rwall[upw], bw_[upw]
        dbwrp = np.where(rWp > 0.0, dbw[upw] / np.where(rWp > 0.0, rWp, 1.0), 0.0)
        bWp = bW[:, upw]
        LvW = Lv[:, IWp]
        lbw = (LvW * bWp).sum(0)
        ccw = ((wv[:, JWp] - v[:, IWp]) * sbar[:, IWp]).sum(0)
        Lvb += _scatter(n, IWp, wsc * bwp * ccw * bWp)
        xbar -= _scatter(n, IWp, ccw * wsc * (dbwrp * lbw * dWp + bwp * _basis_jt(dWp, re, LvW)))
        vbar_in -= _scatter(n, IWp, wsc * bwp * lbw * sbar[:, IWp])
        opbar_v[:, 3, :] = Lvb.T

        # moment operators of both kinds
        for kind_wall, opbar in ((1, opbar_v), (2, opbar_p)):
            xbar += self._moment_adjoint(x, wp, wn, re, quad, wsc, lam_reg, kind_wall, opbar)

        return {"position": xbar, "velocity": vbar_in, "pressure": pbar, "gravity": gbar,
                "lambda": lam, "x_corrected": xpp}

    def _moment_adjoint(self, x, wp, wn, re, quad, wsc, lam_reg, boundary, opbar):
        n = x.shape[1]
        I, J, IW, JW = self.I, self.J, self.IW, self.JW
        d = x[:, J] - x[:, I]
        r = np.sqrt((d * d).sum(0))
        w = np.where(r > 0.0, _weight(r, re, quad), 0.0)
        dw = np.where(r > 0.0, _slope(r, re, quad), 0.0)
        u = w > 0.0
        b = _basis(d[:, u], re)
        outer = ((w[u] * b)[:, None, :] * b[None, :, :]).reshape(81, -1)
        M = _scatter(n, I[u], outer).T.copy()
        dW = wp[:, JW] - x[:, IW]
        rW = np.sqrt((dW * dW).sum(0))
        wW = wsc * np.where(rW > 0.0, _weight(rW, re, quad), 0.0)
        dwW = wsc * np.where(rW > 0.0, _slope(rW, re, quad), 0.0)
        uw = wW > 0.0
        vW = (_neumann(dW[:, uw], wn[:, JW[uw]], re) if boundary == 2 else _basis(dW[:, uw], re))
        outerW = ((wW[uw] * vW)[:, None, :] * vW[None, :, :]).reshape(81, -1)
        M += _scatter(n, IW[uw], outerW).T
        M = M.reshape(n, 9, 9)
        tr = np.trace(M, axis1=1, axis2=2)
        reg = lam_reg * np.where(tr > 0.0, tr / 9.0, 1.0)
        M = M + reg[:, None, None] * np.eye(9)
        minv = np.linalg.inv(M)
        # gradient with respect to the inverse: rows 0..2 and 3+4+5
        Ab = np.zeros((n, 9, 9))
        Ab[:, 0:3, :] = opbar[:, 0:3, :]
        Ab[:, 3:6, :] += opbar[:, 3:4, :]
        mt = minv.transpose(0, 2, 1)
        Mb = -mt @ Ab @ mt
        S = 0.5 * (Mb + Mb.transpose(0, 2, 1))
        ts = np.where(tr > 0.0, np.trace(S, axis1=1, axis2=2), 0.0)
        S = S + (lam_reg / 9.0 * ts)[:, None, None] * np.eye(9)
        xb = np.zeros((3, n))
        Iu = I[u]
        Sb = np.einsum("nkl,ln->kn", S[Iu], b)
        vsv = (b * Sb).sum(0)
        du, ru = d[:, u], r[u]
        gp = dw[u] / ru * vsv * du + 2.0 * w[u] * _basis_jt(du, re, Sb)
        xb += _scatter(n, J[u], gp) - _scatter(n, Iu, gp)
        IWu = IW[uw]
        SvW = np.einsum("nkl,ln->kn", S[IWu], vW)
        vsvW = (vW * SvW).sum(0)
        dWu, rWu = dW[:, uw], rW[uw]
        jt = (_neumann_jt(wn[:, JW[uw]], re, SvW) if boundary == 2 else _basis_jt(dWu, re, SvW))
        gw = dwW[uw] / rWu * vsvW * dWu + 2.0 * wW[uw] * jt
        xb -= _scatter(n, IWu, gw)
        return xb
