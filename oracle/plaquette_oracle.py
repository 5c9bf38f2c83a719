"""CPU FP64 restatement of the reference's 2x2-plaquette energy path -- TEST
INFRASTRUCTURE (imported by tests/ only; the product path never calls it).

Follows, in the reference's index lists and op order:
* ``tout_site``                       src/model/honeycomb.jl:33-38
* ``ising_h4`` / ``hb_h4_ZZ``         src/model/basis.jl:37-55, src/model/honeycomb.jl:46-68
* ``get_local_h``                     src/model/local_h.jl:1-31 (sequential SVDs: a
                                      four-site h as the product s1 s2 s3 s4)
* ``get_op_Ad4``                      src/graph/evaluation.jl:504-519
* ``energy_norm_4x4``                 src/graph/evaluation.jl:430-496
* ``get_gs_energy_4x4``               src/graph/evaluation.jl:340-355

Pinned (tests/test_oracle_plaquette.py): the reference's KAT of this path
(test/test_local_h4.jl: e0 / 8 = -0.16349 +- 1e-5 for the loop-gas state in
the ZZ geometry), the exact reconstruction of h from get_local_h, and the
product-state closed form (D = 1: the plaquette energy is the expectation
of h in the product state).
"""
import math

import numpy as np
import torch

import ctmrg_oracle as O
from ctmrg_oracle import reshape_f, tcon


def tout_site(a, b):
    """c[i, j, k, l] = a[i, j] b[k, l] reshaped ((i, j), (k, l)) (honeycomb.jl:33-38)."""
    c = tcon([a, b], [[-1, -2], [-3, -4]])
    s = c.shape
    return reshape_f(c, s[0] * s[1], s[2] * s[3])


def ising_h4(h=2.5):
    """basis.jl:37-55."""
    SI, sx, sz = O.sI, O.sigmax, O.sigmaz
    II = tout_site(SI, SI)
    Ix = tout_site(SI, sx)
    xI = tout_site(sx, SI)
    zI = h * tout_site(sz, SI)
    xx = tout_site(sx, sx)
    h12, h23, h34, h41 = tout_site(xx, II), tout_site(Ix, xI), tout_site(II, xx), tout_site(xI, Ix)
    hz1, hz2 = tout_site(zI, II), tout_site(II, zI)
    H = -h12 - h23 - h34 - h41 - hz1 - hz2
    d = SI.shape[0]
    return reshape_f(H, d * d, d * d, d * d, d * d) / 2


def hb_h4_ZZ():
    """honeycomb.jl:46-68 (sites 1-2-3-4 around the plaquette)."""
    SI = O.sI.to(O.C128)
    sx, sy, sz = O.sigmax.to(O.C128), O.sigmay, O.sigmaz.to(O.C128)
    II, Iy, yI = O.tout(SI, SI), O.tout(SI, sy), O.tout(sy, SI)
    Ix, xI, zz = O.tout(SI, sx), O.tout(sx, SI), O.tout(sz, sz)
    hx1 = tout_site(tout_site(II, Ix), tout_site(xI, II))
    hx2 = tout_site(tout_site(Ix, II), tout_site(II, xI))
    hy1 = tout_site(tout_site(Iy, yI), tout_site(II, II))
    hy2 = tout_site(tout_site(II, II), tout_site(yI, Iy))
    hz1 = tout_site(tout_site(II, zz), tout_site(II, II))
    hz2 = tout_site(tout_site(II, II), tout_site(II, zz))
    H = -hx1 - hx2 - hy1 - hy2 - hz1 - hz2
    d = II.shape[0]
    return reshape_f(H, d * d, d * d, d * d, d * d) / 2


def _svd_sqrt(x):
    """wrap_svd (cutoff 1e-12, no cap) then u sqrt(s), sqrt(s) v^H."""
    u, s, v = O.wrap_svd(x)
    r = torch.sqrt(s).to(x.dtype)
    return u * r, r[:, None] * v.mH


def get_local_h(h):
    """local_h.jl:1-31: [s1 (d,d,r1), s2 (r1,d,d,r2), s3 (r2,d,d,r3), s4 (r3,d,d)]."""
    with torch.no_grad():
        dims = h.shape
        d = int(round(math.sqrt(dims[0])))
        sl = []
        s1, sd = _svd_sqrt(reshape_f(h, dims[0], -1))
        sl.append(reshape_f(s1, d, d, -1))
        for i in range(len(dims) - 2):
            dl = sd.shape[0]
            si, sd = _svd_sqrt(reshape_f(sd, dl * dims[i], -1))
            sl.append(reshape_f(si, dl, d, d, -1))
        dl = sd.shape[0]
        sl.append(reshape_f(sd, dl, d, d))
        return sl


def get_op_Ad4(s1, s2, s3, s4, A1d, A2d, A3d, A4d):
    """evaluation.jl:504-519."""
    o1 = tcon([s1, A1d], [[-6, 1, -4], [-1, -2, -3, -5, 1]])
    d = o1.shape
    o1 = reshape_f(o1, d[0], d[1], d[2], d[3] * d[4], d[5])
    o2 = tcon([s2, A2d], [[-2, -7, 1, -4], [-1, -3, -5, -6, 1]])
    d = o2.shape
    o2 = reshape_f(o2, d[0], d[1] * d[2], d[3] * d[4], d[5], d[6])
    o3 = tcon([s3, A3d], [[-1, -7, 1, -3], [-2, -4, -5, -6, 1]])
    d = o3.shape
    o3 = reshape_f(o3, d[0] * d[1], d[2] * d[3], d[4], d[5], d[6])
    o4 = tcon([s4, A4d], [[-4, -6, 1], [-1, -2, -3, -5, 1]])
    d = o4.shape
    o4 = reshape_f(o4, d[0], d[1], d[2], d[3] * d[4], d[5])
    return o1, o2, o3, o4


def _mat3(x):
    s = x.shape
    return reshape_f(x, s[0] * s[1] * s[2], -1)


def energy_norm_4x4(C1, C2, C3, C4, E1l, E1r, E2u, E2d, E3l, E3r, E4u, E4d, A1, A2, A3, A4, A1d, A2d, A3d, A4d,
                    oA1d, oA2d, oA3d, oA4d):
    """evaluation.jl:443-496: (energy, norm) of the 2 x 2 plaquette."""
    C1E1 = tcon([C1, E1l], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([C1E1, E4u], [[1, -2, -5, -3], [1, -1, -6, -4]])
    CEEA1 = tcon([CEE4, A1], [[-1, -2, -3, -4, 1, 2], [1, 2, -5, -6, -7]])
    e_LU = _mat3(tcon([CEEA1, oA1d], [[-1, -4, 1, 2, -2, -5, 3], [1, 2, -3, -6, 3]]))
    n_LU = _mat3(tcon([CEEA1, A1d], [[-1, -4, 1, 2, -2, -5, 3], [1, 2, -3, -6, 3]]))

    C4E3 = tcon([C4, E3l], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([C4E3, E4d], [[1, -2, -6, -4], [-1, 1, -5, -3]])
    CEEA4 = tcon([CEE4, A4], [[-1, -2, -3, -4, 1, 2], [-5, 1, 2, -6, -7]])
    e_LB = _mat3(tcon([CEEA4, oA4d], [[-1, -4, 1, 2, -2, -5, 3], [-3, 1, 2, -6, 3]]))
    n_LB = _mat3(tcon([CEEA4, A4d], [[-1, -4, 1, 2, -2, -5, 3], [-3, 1, 2, -6, 3]]))

    C2E1 = tcon([C2, E1r], [[1, -2], [-1, 1, -3, -4]])
    CEE2 = tcon([C2E1, E2u], [[-1, 1, -5, -3], [1, -2, -6, -4]])
    CEEA2 = tcon([CEE2, A2], [[-1, -2, -3, -4, 1, 2], [1, -5, -6, 2, -7]])
    e_RU = _mat3(tcon([CEEA2, oA2d], [[-1, -4, 1, 2, -2, -5, 3], [1, -3, -6, 2, 3]]))
    n_RU = _mat3(tcon([CEEA2, A2d], [[-1, -4, 1, 2, -2, -5, 3], [1, -3, -6, 2, 3]]))

    C3E3 = tcon([C3, E3r], [[-1, 1], [-2, 1, -3, -4]])
    CEE2 = tcon([C3E3, E2d], [[1, -2, -5, -3], [-1, 1, -6, -4]])
    CEEA3 = tcon([CEE2, A3], [[-1, -2, -5, -6, 1, 2], [-3, -4, 1, 2, -7]])
    e_RB = _mat3(tcon([CEEA3, oA3d], [[-1, -4, -2, -5, 1, 2, 3], [-3, -6, 1, 2, 3]]))
    n_RB = _mat3(tcon([CEEA3, A3d], [[-1, -4, -2, -5, 1, 2, 3], [-3, -6, 1, 2, 3]]))

    e_left = tcon([e_LU.transpose(0, 1), e_LB], [[-1, 1], [1, -2]])
    e_right = tcon([e_RU, e_RB], [[-1, 1], [1, -2]])
    energy = tcon([e_left, e_right], [[1, 2], [1, 2]])
    n_left = tcon([n_LU.transpose(0, 1), n_LB], [[-1, 1], [1, -2]])
    n_right = tcon([n_RU, n_RB], [[-1, 1], [1, -2]])
    norm = tcon([n_left, n_right], [[1, 2], [1, 2]])
    return energy, norm


def get_gs_energy_4x4(ts, sl):
    """evaluation.jl:340-355: energy / norm with A1..A4 = A."""
    A, Ad = ts.A, ts.Ad
    s1, s2, s3, s4 = sl
    o = get_op_Ad4(s1, s2, s3, s4, Ad, Ad, Ad, Ad)
    C1, C2, C3, C4 = ts.Cs
    E1, E2, E3, E4 = ts.Es
    e, n = energy_norm_4x4(C1, C2, C3, C4, E1, E1, E2, E2, E3, E3, E4, E4, A, A, A, A, Ad, Ad, Ad, Ad, *o)
    return e / n, e, n
