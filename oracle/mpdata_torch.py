"""TEST INFRASTRUCTURE ONLY: the standard and the infinite-gauge
non-oscillatory MPDATA steps (2 iterations, open walls, ih = 1) written with
vectorised torch operations on the CPU, so that torch.autograd gives an
independent reference gradient for the hand-written adjoint kernels.

The operations and their order follow the oracle (mpdata_oracle_impl.inc),
which is pinned to the reference: upwind3d-gpubc.f90:101-198,
antidiff3d-standard-gpubc.f90:59-462 and antidiff3d-gauge-gpubc.f90:71-430
(with the oracle's ground-plane bound, the 6-point minimum of
mpdata3d-rho-gauge-gpubc.f90:365-368).  Every branch is a torch.where on the
forward value (|y| has derivative 0 at y = 0), the linearisation the
tangent-linear and adjoint kernels use.  Arrays are (k, j, i) with Fortran
index = array index (halo index 0).
"""
import torch

EP = 1.e-10


def _pp(y):
    return torch.where(y > 0, y, torch.zeros_like(y))


def _mn0(y):
    return torch.where(y < 0, y, torch.zeros_like(y))


def _pn(y):
    return -_mn0(y)


def _abs(y):
    return torch.where(y > 0, y, torch.where(y < 0, -y, y * 0))


def _donor(y1, y2, a):
    return _pp(a) * y1 - (-(_mn0(a) * y2))


def _max2(a, b):
    return torch.where(a > b, a, b)


def _min2(a, b):
    return torch.where(a < b, a, b)


def _min3(a, b, c):
    return _min2(_min2(a, b), c)


def _vdyfa(x1, x2, a, r):
    return (_abs(a) - a * a * r) * (_abs(x2) - _abs(x1)) / (_abs(x2) + _abs(x1) + EP)


def _vcorra(a, b, y1, y2, r):
    return -(0.125 * a * b * y1 * r / (y2))


def _corr_terms(c0, c1, c2, c3):
    a0, a1, a2, a3 = _abs(c0), _abs(c1), _abs(c2), _abs(c3)
    return (a0 + a1) - (a2 + a3), a0 + a1 + a2 + a3 + EP


def _sign1(y):
    return torch.copysign(torch.ones_like(y), y).detach()


def _slimv(v, a1, a2, xm, xc):
    sm, smn, sc, scn = _sign1(xm), _sign1(-xm), _sign1(xc), _sign1(-xc)
    return _pp(v) * (a1 * _pp(sm) + a2 * _pp(smn)) - _pn(v) * (a2 * _pp(sc) + a1 * _pp(scn))


def _with_z(base, z):
    """Planes 2..lp-1 take base + z, the ground and sky planes base."""
    return torch.cat([base[:1], base[1:-1] + z, base[-1:]], dim=0)


def _upwind(x, u1, u2, u3, rhr, h, bcx, bcy):
    """The upwind pass: xant with the mirrored x/y halo, and 1/h."""
    lp, mp, np_ = (s - 2 for s in x.shape)
    K, J, I = slice(1, lp + 1), slice(1, mp + 1), slice(1, np_ + 1)
    # putbcinxh (mpdataoperators.inc:253-287)
    xb = x.clone()
    xb[K, J, 0] = bcx[0]
    xb[K, J, np_ + 1] = bcx[1]
    xb[K, 0, I] = bcy[0]
    xb[K, mp + 1, I] = bcy[1]
    # upwind (upwind3d-gpubc.f90:101-198)
    xc = xb[K, J, I]
    F1 = _donor(xb[K, J, 0:np_ + 1], xb[K, J, 1:np_ + 2], u1[K, J, 1:np_ + 2])
    F2 = _donor(xb[K, 0:mp + 1, I], xb[K, 1:mp + 2, I], u2[K, 1:mp + 2, I])
    F3 = _donor(xb[1:lp, J, I], xb[2:lp + 1, J, I], u3[2:lp + 1, J, I])   # faces 2..lp
    f3p = torch.cat([F3, -F3[-1:]], dim=0)
    f3 = torch.cat([-F3[:1], F3], dim=0)
    hinv = 1.0 / h[K, J, I]
    xa = rhr[K, J, I] * (xc - (F1[..., 1:] - F1[..., :-1] + F2[:, 1:, :] - F2[:, :-1, :] +
                               f3p - f3) * hinv)
    # xant with the mirrored x/y halo (cp_scnd_last_to_halo_xy_fulld), y first
    psi = torch.zeros_like(x)
    psi[K, J, I] = xa
    psi[:, 0, :] = psi[:, 2, :]
    psi[:, mp + 1, :] = psi[:, mp - 1, :]
    psi[:, :, 0] = psi[:, :, 2]
    psi[:, :, np_ + 1] = psi[:, :, np_ - 1]
    return psi, hinv


def _limiter(x, P, h, w1, w2, w3):
    """cp, cn (zero halo) of the non-oscillatory limiter
    (antidiff3d-standard-gpubc.f90:260-347, antidiff3d-gauge-gpubc.f90:
    238-324) with the bounds of x (edge-copied halo) and P, and the face
    quantities w on the interior faces (x faces 2..np, y faces 2..mp, z faces
    2..lp; walls zero, lids mirrored)."""
    lp, mp, np_ = (s - 2 for s in x.shape)
    K, J, I = slice(1, lp + 1), slice(1, mp + 1), slice(1, np_ + 1)
    # x with edge-copied halo (cp_last_to_halo_xy), interior ranges
    x0 = x.clone()
    x0[K, 0, I] = x0[K, 1, I]
    x0[K, mp + 1, I] = x0[K, mp, I]
    x0[K, J, 0] = x0[K, J, 1]
    x0[K, J, np_ + 1] = x0[K, J, np_]
    mxo = x0[K, J, 0:np_]
    mno = mxo
    mx = P[K, J, 0:np_]
    mn = mx
    for src, arr in ((0, x0), (1, P)):
        lo, hi = (mno, mxo) if src == 0 else (mn, mx)
        for t in (arr[K, J, I], arr[K, J, 2:np_ + 2], arr[K, 0:mp, I], arr[K, 2:mp + 2, I]):
            lo, hi = _min2(lo, t), _max2(hi, t)
        if src == 0:
            mno, mxo = lo, hi
        else:
            mn, mx = lo, hi
    # k-1 (not on the ground plane), then k+1 (not on the sky plane)
    for dk, sel in ((-1, slice(1, lp)), (1, slice(0, lp - 1))):
        rows = slice(1 + dk + sel.start, 1 + dk + sel.stop)
        for src, arr in ((0, x0), (1, P)):
            t = arr[rows, J, I]
            lo, hi = (mno, mxo) if src == 0 else (mn, mx)
            keep = [slice(0, sel.start), slice(sel.stop, lp)]
            lo = torch.cat([lo[keep[0]], _min2(lo[sel], t), lo[keep[1]]], dim=0)
            hi = torch.cat([hi[keep[0]], _max2(hi[sel], t), hi[keep[1]]], dim=0)
            if src == 0:
                mno, mxo = lo, hi
            else:
                mn, mx = lo, hi
    mx = _max2(mx, mxo)
    mn = _min2(mn, mno)
    zx = torch.zeros_like(w1[..., :1])
    zy = torch.zeros_like(w2[:, :1, :])
    ww1 = torch.cat([zx, w1, zx], dim=2)     # faces 1..np+1, walls 0
    ww2 = torch.cat([zy, w2, zy], dim=1)
    a1p, a1 = ww1[..., 1:], ww1[..., :-1]
    a2p, a2 = ww2[:, 1:, :], ww2[:, :-1, :]
    a3p = torch.cat([w3, -w3[-1:]], dim=0)
    a3 = torch.cat([-w3[:1], w3], dim=0)
    xcc, hc = P[K, J, I], h[K, J, I]
    cpi = (mx - xcc) * hc / (_pn(a1p) + _pp(a1) + _pn(a2p) + _pp(a2) + _pn(a3p) + _pp(a3) + EP)
    cni = (xcc - mn) * hc / (_pp(a1p) + _pn(a1) + _pp(a2p) + _pn(a2) + _pp(a3p) + _pn(a3) + EP)
    cp = torch.zeros_like(x)
    cn = torch.zeros_like(x)
    cp[K, J, I] = cpi
    cn[K, J, I] = cni
    return cp, cn


def step(x, u1, u2, u3, rhr, h, bcx, bcy):
    """One standard step; returns the new field (interior; halo copied
    from x)."""
    lp, mp, np_ = (s - 2 for s in x.shape)
    K, J, I = slice(1, lp + 1), slice(1, mp + 1), slice(1, np_ + 1)
    psi, hinv = _upwind(x, u1, u2, u3, rhr, h, bcx, bcy)
    # pseudo-velocities (antidiff3d-standard-gpubc.f90:72-202), interior faces
    P = psi
    # x faces i = 2..np
    hm = 1.0 / (0.5 * (h[K, J, 1:np_] + h[K, J, 2:np_ + 1]))
    a = u1[K, J, 2:np_ + 1]
    s2 = u2[K, 1:mp + 1, 1:np_] + u2[K, 2:mp + 2, 1:np_] + u2[K, 2:mp + 2, 2:np_ + 1] + \
        u2[K, 1:mp + 1, 2:np_ + 1]
    y1, y2 = _corr_terms(P[K, 2:mp + 2, 1:np_], P[K, 2:mp + 2, 2:np_ + 1], P[K, 0:mp, 1:np_],
                         P[K, 0:mp, 2:np_ + 1])
    base = _vdyfa(P[K, J, 1:np_], P[K, J, 2:np_ + 1], a, hm) + _vcorra(a, s2, y1, y2, hm)
    Kz = slice(2, lp)
    s3 = u3[Kz, J, 1:np_] + u3[3:lp + 1, J, 1:np_] + u3[3:lp + 1, J, 2:np_ + 1] + \
        u3[Kz, J, 2:np_ + 1]
    y1, y2 = _corr_terms(P[3:lp + 1, J, 1:np_], P[3:lp + 1, J, 2:np_ + 1], P[1:lp - 1, J, 1:np_],
                         P[1:lp - 1, J, 2:np_ + 1])
    v1 = _with_z(base, _vcorra(a[1:-1], s3, y1, y2, hm[1:-1]))
    fns1 = _donor(P[K, J, 1:np_], P[K, J, 2:np_ + 1], v1)
    # y faces j = 2..mp
    hm = 1.0 / (0.5 * (h[K, 1:mp, I] + h[K, 2:mp + 1, I]))
    a = u2[K, 2:mp + 1, I]
    s1 = u1[K, 1:mp, 1:np_ + 1] + u1[K, 2:mp + 1, 1:np_ + 1] + u1[K, 2:mp + 1, 2:np_ + 2] + \
        u1[K, 1:mp, 2:np_ + 2]
    y1, y2 = _corr_terms(P[K, 1:mp, 2:np_ + 2], P[K, 2:mp + 1, 2:np_ + 2], P[K, 1:mp, 0:np_],
                         P[K, 2:mp + 1, 0:np_])
    base = _vdyfa(P[K, 1:mp, I], P[K, 2:mp + 1, I], a, hm) + _vcorra(a, s1, y1, y2, hm)
    s3 = u3[Kz, 1:mp, I] + u3[Kz, 2:mp + 1, I] + u3[3:lp + 1, 2:mp + 1, I] + u3[3:lp + 1, 1:mp, I]
    y1, y2 = _corr_terms(P[3:lp + 1, 1:mp, I], P[3:lp + 1, 2:mp + 1, I], P[1:lp - 1, 1:mp, I],
                         P[1:lp - 1, 2:mp + 1, I])
    v2 = _with_z(base, _vcorra(a[1:-1], s3, y1, y2, hm[1:-1]))
    fns2 = _donor(P[K, 1:mp, I], P[K, 2:mp + 1, I], v2)
    # z faces k = 2..lp
    hm = 1.0 / (0.5 * (h[1:lp, J, I] + h[2:lp + 1, J, I]))
    a = u3[2:lp + 1, J, I]
    s1 = u1[1:lp, J, 1:np_ + 1] + u1[2:lp + 1, J, 1:np_ + 1] + u1[2:lp + 1, J, 2:np_ + 2] + \
        u1[1:lp, J, 2:np_ + 2]
    s2 = u2[1:lp, 1:mp + 1, I] + u2[1:lp, 2:mp + 2, I] + u2[2:lp + 1, 2:mp + 2, I] + \
        u2[2:lp + 1, 1:mp + 1, I]
    y1a, y2a = _corr_terms(P[1:lp, J, 2:np_ + 2], P[2:lp + 1, J, 2:np_ + 2], P[1:lp, J, 0:np_],
                           P[2:lp + 1, J, 0:np_])
    y1b, y2b = _corr_terms(P[1:lp, 2:mp + 2, I], P[2:lp + 1, 2:mp + 2, I], P[1:lp, 0:mp, I],
                           P[2:lp + 1, 0:mp, I])
    v3 = _vdyfa(P[1:lp, J, I], P[2:lp + 1, J, I], a, hm) + _vcorra(a, s1, y1a, y2a, hm) + \
        _vcorra(a, s2, y1b, y2b, hm)
    fns3 = _donor(P[1:lp, J, I], P[2:lp + 1, J, I], v3)
    # limiter (antidiff3d-standard-gpubc.f90:260-347) on the donor fluxes of V
    cp, cn = _limiter(x, P, h, fns1, fns2, fns3)
    zx = torch.zeros_like(fns1[..., :1])
    zy = torch.zeros_like(fns2[:, :1, :])
    # limited donor fluxes (:372-407), interior faces
    one = torch.ones((), dtype=x.dtype)
    rc = P[K, J, I] / rhr[K, J, I]

    def flux(v, pm, pc, rm, rcv, cpc, cnm, cpm, cnc):
        vl = _slimv(v, _min3(one, cpc, cnm), _min3(one, cpm, cnc), pm, pc)
        return _donor(pm / rm, rcv, vl)

    f1 = flux(v1, P[K, J, 1:np_], P[K, J, 2:np_ + 1], rhr[K, J, 1:np_], rc[..., 1:],
              cp[K, J, 2:np_ + 1], cn[K, J, 1:np_], cp[K, J, 1:np_], cn[K, J, 2:np_ + 1])
    f2 = flux(v2, P[K, 1:mp, I], P[K, 2:mp + 1, I], rhr[K, 1:mp, I], rc[:, 1:, :],
              cp[K, 2:mp + 1, I], cn[K, 1:mp, I], cp[K, 1:mp, I], cn[K, 2:mp + 1, I])
    f3i = flux(v3, P[1:lp, J, I], P[2:lp + 1, J, I], rhr[1:lp, J, I], rc[1:],
               cp[2:lp + 1, J, I], cn[1:lp, J, I], cp[1:lp, J, I], cn[2:lp + 1, J, I])
    # update (:431-459)
    g1 = torch.cat([zx, f1, zx], dim=2)
    g2 = torch.cat([zy, f2, zy], dim=1)
    s = g1[..., 1:] - g1[..., :-1] + g2[:, 1:, :] - g2[:, :-1, :]
    first = s[:1] + f3i[:1] + f3i[:1]
    mid = s[1:-1] + f3i[1:] - f3i[:-1]
    last = s[-1:] - f3i[-1:] - f3i[-1:]
    ssum = torch.cat([first, mid, last], dim=0)
    out = x.clone()
    out[K, J, I] = rhr[K, J, I] * (P[K, J, I] / rhr[K, J, I] - ssum * hinv)
    return out


def _rat4(z0, z1, z2, z3):
    return ((z3 + z2) - (z1 + z0)) * 0.25


def _vdyf(x1, x2, a, r):
    return (_abs(a) - a * a * r) * ((x2 - x1) * 0.5)


def _gflux(v, cpc, cnm, cpm, cnc):
    one = torch.ones((), dtype=v.dtype)
    return _pp(v) * _min3(one, cpc, cnm) - _pn(v) * _min3(one, cpm, cnc)


def step_gauge(x, u1, u2, u3, rhr, h, bcx, bcy):
    """One infinite-gauge step (antidiff3d-gauge-gpubc.f90:71-430); returns
    the new field (interior; halo copied from x)."""
    lp, mp, np_ = (s - 2 for s in x.shape)
    K, J, I = slice(1, lp + 1), slice(1, mp + 1), slice(1, np_ + 1)
    Kz = slice(2, lp)
    psi, _ = _upwind(x, u1, u2, u3, rhr, h, bcx, bcy)
    P = psi
    # pseudo-velocities (:89-203), interior faces
    # x faces i = 2..np
    hm = 1.0 / (0.5 * (h[K, J, 1:np_] + h[K, J, 2:np_ + 1]))
    a = u1[K, J, 2:np_ + 1]
    s2 = u2[K, 1:mp + 1, 1:np_] + u2[K, 2:mp + 2, 1:np_] + u2[K, 2:mp + 2, 2:np_ + 1] + \
        u2[K, 1:mp + 1, 2:np_ + 1]
    cr = s2 * _rat4(P[K, 0:mp, 1:np_], P[K, 0:mp, 2:np_ + 1], P[K, 2:mp + 2, 1:np_],
                    P[K, 2:mp + 2, 2:np_ + 1])
    s3 = u3[Kz, J, 1:np_] + u3[3:lp + 1, J, 1:np_] + u3[3:lp + 1, J, 2:np_ + 1] + \
        u3[Kz, J, 2:np_ + 1]
    cr = _with_z(cr, s3 * _rat4(P[1:lp - 1, J, 1:np_], P[1:lp - 1, J, 2:np_ + 1],
                                P[3:lp + 1, J, 1:np_], P[3:lp + 1, J, 2:np_ + 1]))
    v1 = _vdyf(P[K, J, 1:np_], P[K, J, 2:np_ + 1], a, hm) - 0.125 * a * hm * (cr)
    # y faces j = 2..mp
    hm = 1.0 / (0.5 * (h[K, 1:mp, I] + h[K, 2:mp + 1, I]))
    a = u2[K, 2:mp + 1, I]
    s1 = u1[K, 1:mp, 1:np_ + 1] + u1[K, 2:mp + 1, 1:np_ + 1] + u1[K, 2:mp + 1, 2:np_ + 2] + \
        u1[K, 1:mp, 2:np_ + 2]
    cr = s1 * _rat4(P[K, 1:mp, 0:np_], P[K, 2:mp + 1, 0:np_], P[K, 1:mp, 2:np_ + 2],
                    P[K, 2:mp + 1, 2:np_ + 2])
    s3 = u3[Kz, 1:mp, I] + u3[Kz, 2:mp + 1, I] + u3[3:lp + 1, 2:mp + 1, I] + u3[3:lp + 1, 1:mp, I]
    cr = _with_z(cr, s3 * _rat4(P[1:lp - 1, 1:mp, I], P[1:lp - 1, 2:mp + 1, I],
                                P[3:lp + 1, 1:mp, I], P[3:lp + 1, 2:mp + 1, I]))
    v2 = _vdyf(P[K, 1:mp, I], P[K, 2:mp + 1, I], a, hm) - 0.125 * a * hm * (cr)
    # z faces k = 2..lp
    hm = 1.0 / (0.5 * (h[1:lp, J, I] + h[2:lp + 1, J, I]))
    a = u3[2:lp + 1, J, I]
    s1 = u1[1:lp, J, 1:np_ + 1] + u1[2:lp + 1, J, 1:np_ + 1] + u1[2:lp + 1, J, 2:np_ + 2] + \
        u1[1:lp, J, 2:np_ + 2]
    s2 = u2[1:lp, 1:mp + 1, I] + u2[1:lp, 2:mp + 2, I] + u2[2:lp + 1, 2:mp + 2, I] + \
        u2[2:lp + 1, 1:mp + 1, I]
    cr = s1 * _rat4(P[1:lp, J, 0:np_], P[2:lp + 1, J, 0:np_], P[1:lp, J, 2:np_ + 2],
                    P[2:lp + 1, J, 2:np_ + 2]) + \
        s2 * _rat4(P[1:lp, 0:mp, I], P[2:lp + 1, 0:mp, I], P[1:lp, 2:mp + 2, I],
                   P[2:lp + 1, 2:mp + 2, I])
    v3 = _vdyf(P[1:lp, J, I], P[2:lp + 1, J, I], a, hm) - 0.125 * a * hm * (cr)
    # limiter (:238-324) on the pseudo-velocities
    cp, cn = _limiter(x, P, h, v1, v2, v3)
    # limited fluxes and update (:340-428): faces 1..n+1 (wall V zero), the
    # lid faces from the mirrored V
    zx = torch.zeros_like(v1[..., :1])
    zy = torch.zeros_like(v2[:, :1, :])
    V1 = torch.cat([zx, v1, zx], dim=2)
    V2 = torch.cat([zy, v2, zy], dim=1)
    F1 = _gflux(V1, cp[K, J, 1:np_ + 2], cn[K, J, 0:np_ + 1], cp[K, J, 0:np_ + 1],
                cn[K, J, 1:np_ + 2])
    F2 = _gflux(V2, cp[K, 1:mp + 2, I], cn[K, 0:mp + 1, I], cp[K, 0:mp + 1, I],
                cn[K, 1:mp + 2, I])
    Fi = _gflux(v3, cp[2:lp + 1, J, I], cn[1:lp, J, I], cp[1:lp, J, I], cn[2:lp + 1, J, I])
    Fg = _gflux(-v3[:1], cp[1:2, J, I], cn[2:3, J, I], cp[2:3, J, I], cn[1:2, J, I])
    Fs = _gflux(-v3[-1:], cp[lp - 1:lp, J, I], cn[lp:lp + 1, J, I], cp[lp:lp + 1, J, I],
                cn[lp - 1:lp, J, I])
    F3 = torch.cat([Fg, Fi, Fs], dim=0)      # faces 1..lp+1
    ssum = F1[..., 1:] - F1[..., :-1] + F2[:, 1:, :] - F2[:, :-1, :] + F3[1:] - F3[:-1]
    out = x.clone()
    out[K, J, I] = rhr[K, J, I] * (P[K, J, I] / rhr[K, J, I] - ssum / h[K, J, I])
    return out
