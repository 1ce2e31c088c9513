"""FP64 torch restatement of the differentiable electron run of
include/gasx.h (gasx_diff_*), for tests: autograd through it is the
reference gradient of the hand-written adjoint.

Same model as the library: prologue half drift + wedge map (EM2.cu:14-56),
then per step the field cycle with the current of the previous deposit
(EM3_1.cu:9-15, EM3.cu:14-42, cycle of electricmagneticfield6.m:146-151),
gather + relativistic push (EM4.cu:36-148), half drift + wedge map, the
recorded collision kick and likelihood-ratio weight, half drift + wedge map,
and the area-weighted deposit (EM7.cu:56-202) scaled by the amplitude.  The
objective is PHIp eps0 / 2 sum Er^2 r dr dz over the Er cells with
(i + 1/2) dr >= r_frac Rp at the last step.

Arrays are indexed [k, i] (z rows, r columns), which is the reference's
i + k*nr layout read as a (rows, cols) matrix.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

T = torch.float64


@dataclass
class Grid:
    nr: int
    nz: int
    dr: float
    dz: float
    phip: float
    ddt: float
    n_sub: int
    fn: float
    c: float
    eps0: float
    mu0: float
    e_em: float
    ez_metric: int = 0
    axis_sign: int = 1

    @classmethod
    def from_params(cls, p):
        return cls(p.nr, p.nz, p.dr, p.dz, p.phip, p.ddt, p.n_sub, p.fn, p.c, p.eps0, p.mu0, p.e_em,
                   p.ez_metric, 1 if p.axis_j_sign >= 0 else -1)

    @property
    def rp(self):
        return self.nr * self.dr

    @property
    def lp(self):
        return self.nz * self.dz


# ---- field cycle ------------------------------------------------------------------
def h_update(g: Grid, s, er, ez, h):
    dtf = g.ddt / g.n_sub
    ch_r, ch_z = dtf / (g.mu0 * g.dr), dtf / (g.mu0 * g.dz)
    return h + s * ch_r * (ez[:, 1:] - ez[:, :-1]) - s * ch_z * (er[1:, :] - er[:-1, :])


def e_update(g: Grid, er, ez, h, jr, jz):
    dtf = g.ddt / g.n_sub
    ce, cez, c4 = dtf / g.eps0, dtf / (g.eps0 * g.dz), 4 * dtf / (g.eps0 * g.dr)
    inner = er[1:-1] - ce * (jr[:-1] + jr[1:]) / 2 - cez * (h[1:] - h[:-1])
    er_new = torch.cat([er[1:2], inner, er[-2:-1]], 0)
    i = torch.arange(1, g.nr, dtype=T)
    rs = (i + 1 if g.ez_metric else i) * g.dr
    ca = ce * (1 / (2 * rs) + 1 / g.dr)
    cb = ce * (1 / (2 * rs) - 1 / g.dr)
    col0 = ez[:, :1] + g.axis_sign * ce * jz[:, :1] + c4 * h[:, :1]
    mid = ez[:, 1:g.nr] - ce * (jz[:, :-1] + jz[:, 1:]) / 2 + ca * h[:, 1:] + cb * h[:, :-1]
    ez_new = torch.cat([col0, mid, ez[:, g.nr:]], 1)
    return er_new, ez_new


def field_cycle(g: Grid, er, ez, h, jr, jz):
    h = h_update(g, -0.5, er, ez, h)
    for _ in range(g.n_sub):
        h = h_update(g, 1.0, er, ez, h)
        er, ez = e_update(g, er, ez, h, jr, jz)
    h = h_update(g, 0.5, er, ez, h)
    return er, ez, h


# ---- particles ---------------------------------------------------------------------
def half_w(s, n):
    a = torch.floor(s.detach())
    lo, hi = a < 0, a >= n - 1
    edge = lo | hi
    f = s - a
    i0 = torch.where(lo, 0, torch.where(hi, n - 1, a)).long()
    i1 = torch.where(lo, 0, torch.where(hi, n - 1, a + 1)).long()
    w1 = torch.where(edge, torch.zeros_like(f), f)
    return i0, i1, 1 - w1, w1


def node_w(s, n):
    a = torch.floor(s.detach())
    out = (a < 0) | (a >= n)
    f = torch.where(out, torch.zeros_like(s), s - a)
    ac = a.clamp(0, n)
    return ac.long(), (ac + 1).clamp(max=n).long(), 1 - f, f


def wedge(g: Grid, x, y, z, vx, vy, vz):
    """EM2.cu:29-56: returns the mapped state and the absorbed mask."""
    r2 = (x * x + y * y).detach()
    dead = r2 > g.rp * g.rp
    xd, yd = x.detach(), y.detach()
    ang = torch.where(xd > 0, torch.atan2(yd, xd), torch.atan(yd / xd))
    ang = torch.nan_to_num(ang, nan=0.0)
    k_up = torch.ceil(-ang / g.phip)
    k_dn = torch.ceil((ang - g.phip) / g.phip)
    rot = torch.where(ang < 0, k_up * g.phip, torch.where(ang > g.phip, -k_dn * g.phip, torch.zeros_like(ang)))
    sgn = torch.where(xd < 0, -1.0, 1.0).to(T)
    cr, sr = torch.cos(rot), torch.sin(rot)
    x1 = sgn * (cr * x - sr * y)
    y1 = sgn * (sr * x + cr * y)
    vx1 = cr * vx - sr * vy
    vy1 = sr * vx + cr * vy
    top = (z > g.lp).detach()
    z = torch.where(top, 2 * g.lp - z, z)
    vz = torch.where(top, -vz, vz)
    bot = (z < 0).detach()
    z = torch.where(bot, -z, z)
    vz = torch.where(bot, -vz, vz)
    return x1, y1, z, vx1, vy1, vz, dead


def half_drift(g: Grid, x, y, z, vx, vy, vz):
    h = 0.5 * g.ddt
    return x + h * vx, y + h * vy, z + h * vz


def gather(f, ir0, ir1, wr0, wr1, iz0, iz1, wz0, wz1):
    return (wz0 * (wr0 * f[iz0, ir0] + wr1 * f[iz0, ir1]) +
            wz1 * (wr0 * f[iz1, ir0] + wr1 * f[iz1, ir1]))


def push(g: Grid, qm, er, ez, h, x, y, z, vx, vy, vz):
    r = torch.sqrt(x * x + y * y)
    pos = r.detach() > 0
    rs = torch.where(pos, r, torch.ones_like(r))
    cph = torch.where(pos, x / rs, torch.ones_like(r))
    sph = torch.where(pos, y / rs, torch.zeros_like(r))
    vr = vx * cph + vy * sph
    srn, szn = r / g.dr, z / g.dz
    a = half_w(srn - 0.5, g.nr)
    b = node_w(szn, g.nz)
    c = half_w(szn - 0.5, g.nz)
    d = node_w(srn, g.nr)
    E_r = gather(er, a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3])
    H = gather(h, a[0], a[1], a[2], a[3], c[0], c[1], c[2], c[3])
    E_z = gather(ez, d[0], d[1], d[2], d[3], c[0], c[1], c[2], c[3])
    fr = qm * (E_r - vz * g.mu0 * H)
    fz = qm * (E_z + vr * g.mu0 * H)
    v2 = vx * vx + vy * vy + vz * vz
    gm = 1 / torch.sqrt(1 - v2 / g.c ** 2)
    ux = gm * vx + fr * cph * g.ddt
    uy = gm * vy + fr * sph * g.ddt
    uz = gm * vz + fz * g.ddt
    igp = 1 / torch.sqrt(1 + (ux * ux + uy * uy + uz * uz) / g.c ** 2)
    return ux * igp, uy * igp, uz * igp


def deposit(g: Grid, K, x, y, z, vx, vy, vz, w, alive):
    r = torch.sqrt(x * x + y * y)
    pos = r.detach() > 0
    rs = torch.where(pos, r, torch.ones_like(r))
    vr = torch.where(pos, (vx * x + vy * y) / rs, vx)
    ir0, ir1, wr0, wr1 = half_w(r / g.dr - 0.5, g.nr)
    iz0, iz1, wz0, wz1 = half_w(z / g.dz - 0.5, g.nz)
    vol = g.phip / 2 * g.dr * g.dr * g.dz * (2 * torch.arange(g.nr, dtype=T) + 1)
    s = K * w * alive
    jr = torch.zeros(g.nz * g.nr, dtype=T)
    jz = torch.zeros(g.nz * g.nr, dtype=T)
    for ii, kk, ww in ((ir0, iz0, wr0 * wz0), (ir1, iz0, wr1 * wz0), (ir0, iz1, wr0 * wz1), (ir1, iz1, wr1 * wz1)):
        q = s * ww / vol[ii]
        idx = kk * g.nr + ii
        jr = jr.index_add(0, idx, q * vr)
        jz = jz.index_add(0, idx, q * vz)
    return jr.view(g.nz, g.nr), jz.view(g.nz, g.nr)


def profile(nodes, r, rp):
    nn = nodes.numel()
    s = r * nn / rp - 0.5
    sd = s.detach()
    j0 = torch.floor(sd).clamp(0, nn - 2).long()
    f = s - j0.to(T)
    v = nodes[j0] + f * (nodes[j0 + 1] - nodes[j0])
    v = torch.where(sd <= 0, nodes[0].expand_as(v), v)
    return torch.where(sd >= nn - 1, nodes[nn - 1].expand_as(v), v)


def objective(g: Grid, er, r_frac):
    i = torch.arange(g.nr, dtype=T)
    rc = (i + 0.5) * g.dr
    mask = ((i + 0.5) >= r_frac * g.nr).to(T)
    return 0.5 * g.phip * g.eps0 * g.dr * g.dz * (er * er * rc * mask).sum()


def run(g: Grid, state, events, nodes, nodes0, amplitude, n_steps, r_frac, qnum=-1.0,
        mass=9.1096e-31):
    """W of the model.  state: (6, n) initial particles; events: per step a
    dict with dv (3, n), p0 (n), col (n); nodes, amplitude: torch FP64
    (may require grad); nodes0: the recording profile."""
    x, y, z, vx, vy, vz = (torch.as_tensor(np.asarray(state[i], dtype=np.float64)) for i in range(6))
    n = x.numel()
    w = torch.ones(n, dtype=T)
    alive = torch.ones(n, dtype=T)
    nodes0 = torch.as_tensor(np.asarray(nodes0, dtype=np.float64))
    K = amplitude * g.fn * g.e_em * qnum
    qm = qnum * g.e_em / mass
    er = torch.zeros(g.nz + 1, g.nr, dtype=T)
    ez = torch.zeros(g.nz, g.nr + 1, dtype=T)
    h = torch.zeros(g.nz, g.nr, dtype=T)

    def bc(x, y, z, vx, vy, vz, alive):
        # absorbed particles park at rest inside the domain (away from r = 0,
        # where sqrt has no derivative); every contribution of theirs is
        # masked by `alive`
        x, y, z, vx, vy, vz, dead = wedge(g, x, y, z, vx, vy, vz)
        alive = alive * (~dead).to(T)
        keep = alive > 0
        zero = torch.zeros_like(x)
        px, pz = torch.full_like(x, g.rp / 2), torch.full_like(x, g.lp / 2)
        return (torch.where(keep, x, px), torch.where(keep, y, zero), torch.where(keep, z, pz),
                torch.where(keep, vx, zero), torch.where(keep, vy, zero), torch.where(keep, vz, zero), alive)

    x, y, z = half_drift(g, x, y, z, vx, vy, vz)
    x, y, z, vx, vy, vz, alive = bc(x, y, z, vx, vy, vz, alive)
    jr, jz = deposit(g, K, x, y, z, vx, vy, vz, w, alive)
    for s in range(n_steps):
        er, ez, h = field_cycle(g, er, ez, h, jr, jz)
        vx, vy, vz = push(g, qm, er, ez, h, x, y, z, vx, vy, vz)
        x, y, z = half_drift(g, x, y, z, vx, vy, vz)
        x, y, z, vx, vy, vz, alive = bc(x, y, z, vx, vy, vz, alive)
        ev = events[s]
        dv = torch.as_tensor(np.asarray(ev["dv"], dtype=np.float64))
        p0 = torch.as_tensor(np.asarray(ev["p0"], dtype=np.float64))
        col = torch.as_tensor(np.asarray(ev["col"]) != 0)
        vx, vy, vz = vx + dv[0] * alive, vy + dv[1] * alive, vz + dv[2] * alive
        r = torch.sqrt(x * x + y * y)
        n0r = profile(nodes0, r.detach(), g.rp)
        nr_ = profile(nodes, r, g.rp)
        q = nr_ / n0r
        p0c = torch.where(p0 < 1, p0, torch.zeros_like(p0))
        L = torch.where(col, q, (1 - p0c * q) / (1 - p0c))
        L = torch.where((p0 < 1) & (n0r > 0) & (alive > 0), L, torch.ones_like(L))
        w = w * L
        x, y, z = half_drift(g, x, y, z, vx, vy, vz)
        x, y, z, vx, vy, vz, alive = bc(x, y, z, vx, vy, vz, alive)
        if s + 1 < n_steps:
            jr, jz = deposit(g, K, x, y, z, vx, vy, vz, w, alive)
    return objective(g, er, r_frac)


def layout_fields(g: Grid, er, ez, h):
    """Flat reference-layout arrays -> [k, i] matrices."""
    return (torch.as_tensor(np.asarray(er, dtype=np.float64)).view(g.nz + 1, g.nr),
            torch.as_tensor(np.asarray(ez, dtype=np.float64)).view(g.nz, g.nr + 1),
            torch.as_tensor(np.asarray(h, dtype=np.float64)).view(g.nz, g.nr))


def physical_profile(nodes, rp):
    """Node radii of the profile: (j + 1/2) Rp / n."""
    n = len(nodes)
    return (np.arange(n) + 0.5) * rp / n


__all__ = ["Grid", "field_cycle", "push", "deposit", "wedge", "run", "profile", "objective",
           "layout_fields", "physical_profile", "math"]
