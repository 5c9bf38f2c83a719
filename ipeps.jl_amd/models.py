"""Model Hamiltonians and the deterministic loop-gas state (host-side inputs).

* ``tout``        src/model/basis.jl:16-21 (site-1 index fastest)
* ``ising``       src/model/basis.jl:23-29
* ``heisenberg``  src/model/basis.jl:31-35
* ``honeycomb``   src/model/honeycomb.jl:1-14
* ``init_hb_gs``  src/model/lee_gs.jl:2-28 (D = 2 loop-gas state)
* ``act_Q_op``    src/model/lee_gs.jl:54-73 (the Q operator on a site tensor,
                  the "gauge" tangent basis of optim_es)
* ``tout_site``, ``ising_h4``, ``honeycomb_h4``, ``hb_h4_ZZ``
                  src/model/honeycomb.jl:17-68, basis.jl:37-55 (four-site
                  plaquette Hamiltonians)
* ``get_local_h`` src/model/local_h.jl:1-31 (a plaquette Hamiltonian as the
                  product of four local factors, by sequential SVDs)
"""
import math

import numpy as np

sigmax = np.array([[0, 1], [1, 0]], dtype=np.float64)
sigmay = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
sigmaz = np.array([[1, 0], [0, -1]], dtype=np.float64)
sI = np.eye(2)
Sx, Sy, Sz = sigmax / 2, sigmay / 2, sigmaz / 2


def tout(a, b):
    # c[i1, i2, j1, j2] = a[i1, j1] b[i2, j2], reshaped column-major
    c = np.einsum("ac,bd->abcd", a, b)
    s = c.shape
    return c.reshape(s[0] * s[1], s[2] * s[3], order="F")


def ising(h=2.0):
    H = -tout(sigmax, sigmax) - h * tout(sigmaz, sI) / 4 - h * tout(sI, sigmaz) / 4
    return [H, H]


def heisenberg(Jz=1.0):
    H = Jz * tout(Sz, Sz) - tout(Sx, Sx) - tout(Sy, Sy)
    return [H.astype(np.complex128), H.astype(np.complex128)]


def honeycomb(Jx=1.0, Jy=1.0, Jz=1.0, dir="ZZ"):
    def t4(a, b, c, d):
        return tout(tout(a, b), tout(c, d))

    if dir == "ZZ":
        hv = Jx * t4(sI, sigmax, sigmax, sI) + (t4(sigmaz, sigmaz, sI, sI) + t4(sI, sI, sigmaz, sigmaz)) / 4
        hh = Jy * t4(sI, sigmay, sigmay, sI) + (t4(sigmaz, sigmaz, sI, sI) + t4(sI, sI, sigmaz, sigmaz)) / 4
    elif dir == "XX":
        hv = Jy * t4(sI, sigmay, sigmay, sI) + (t4(sigmax, sigmax, sI, sI) + t4(sI, sI, sigmax, sigmax)) / 4
        hh = Jz * t4(sI, sigmaz, sigmaz, sI) + (t4(sigmax, sigmax, sI, sI) + t4(sI, sI, sigmax, sigmax)) / 4
    elif dir == "YY":
        hv = Jz * t4(sI, sigmaz, sigmaz, sI) + (t4(sigmay, sigmay, sI, sI) + t4(sI, sI, sigmay, sigmay)) / 4
        hh = Jx * t4(sI, sigmax, sigmax, sI) + (t4(sigmay, sigmay, sI, sI) + t4(sI, sI, sigmay, sigmay)) / 4
    else:
        raise ValueError(dir)
    return [-np.real(hh), -np.real(hv)]


def get_Q_op():
    Q = np.zeros((2, 2, 2, 2, 2), dtype=np.complex128)
    Q[0, 0, 0] = sI
    Q[0, 1, 1] = sigmax
    Q[1, 0, 1] = sigmay
    Q[1, 1, 0] = sigmaz
    return Q


def act_Q_op(A0, add=0):
    """Q (x) Q applied to a site tensor A0 (D0,D0,D0,D0,4) -> (2 D0, ..., 4),
    with a flux sigma^x / sigma^y / sigma^z inserted on the first Q for
    add = 1 / 2 / 3 (src/model/lee_gs.jl:54-73)."""
    Q = get_Q_op()
    if add == 1:
        Q1 = np.einsum("xbcpq,xa->abcpq", Q, sigmax)
    elif add == 2:
        Q1 = np.einsum("axcpq,xb->abcpq", Q, sigmay)
    elif add == 3:
        Q1 = np.einsum("abxpq,xc->abcpq", Q, sigmaz)
    elif add == 0:
        Q1 = Q
    else:
        raise ValueError(add)
    QQ = np.einsum("abxeg,cdxfh->abcdefgh", Q1, Q)
    QQ = QQ.reshape(2, 2, 2, 2, 4, 4, order="F")
    A = np.einsum("acegix,bdfhx->abcdefghi", QQ, np.asarray(A0))
    D1, D2 = A.shape[0], A.shape[1]
    return A.reshape(D1 * D2, D1 * D2, D1 * D2, D1 * D2, A.shape[8], order="F")


def init_hb_gs(D=2, dir="XX"):
    if D != 2:
        raise NotImplementedError("only the D=2 loop-gas state")
    Q = get_Q_op()
    u = 1 / math.sqrt(3)
    s111 = np.array([1 + u, u + 1j * u]) / math.sqrt(2 + 2 * u)
    T = np.einsum("abcdp,p->abcd", Q, s111)
    if dir == "ZZ":
        A = np.einsum("abxe,cdxf->abcdef", T, T)
    elif dir == "XX":
        A = np.einsum("xabe,xcdf->abcdef", T, T)
    elif dir == "YY":
        A = np.einsum("bxae,dxcf->abcdef", T, T)
    else:
        raise ValueError(dir)
    return A.reshape(2, 2, 2, 2, 4, order="F")


def seeded_tensor(shape, seed, unit_max=False, dtype=np.float64):
    """numpy default_rng(seed).standard_normal in Fortran order."""
    v = np.random.default_rng(seed).standard_normal(int(np.prod(shape)))
    a = v.reshape(shape, order="F")
    if unit_max:
        a = a / np.abs(a).max()
    return np.asfortranarray(a.astype(dtype))


def tout_site(a, b):
    """c[(i, j), (k, l)] = a[i, j] b[k, l] (honeycomb.jl:33-38)."""
    c = np.einsum("ij,kl->ijkl", a, b)
    sh = c.shape
    return c.reshape(sh[0] * sh[1], sh[2] * sh[3], order="F")


def ising_h4(h=2.5):
    """Transverse-field Ising on a 2 x 2 plaquette (basis.jl:37-55)."""
    II, Ix, xI = tout_site(sI, sI), tout_site(sI, sigmax), tout_site(sigmax, sI)
    zI, xx = h * tout_site(sigmaz, sI), tout_site(sigmax, sigmax)
    H = -tout_site(xx, II) - tout_site(Ix, xI) - tout_site(II, xx) - tout_site(xI, Ix) - tout_site(zI, II) \
        - tout_site(II, zI)
    d = sI.shape[0]
    return H.reshape(d * d, d * d, d * d, d * d, order="F") / 2


def honeycomb_h4():
    """The XX-geometry honeycomb plaquette (honeycomb.jl:17-31)."""
    II, Iy, yI = tout(sI, sI), tout(sI, sigmay), tout(sigmay, sI)
    Iz, zI, xx = tout(sI, sigmaz), tout(sigmaz, sI), tout(sigmax, sigmax)
    h1 = tout(tout(Iy, yI), tout(II, II))
    h2 = tout(tout(II, II), tout(Iy, yI))
    h3 = tout(tout(Iz, II), tout(zI, II))
    h4 = tout(tout(II, Iz), tout(II, zI))
    h5 = tout(tout(II, xx), tout(II, II))
    h6 = tout(tout(II, II), tout(xx, II))
    return (-h1 - h3 - h5 - h2 - h4 - h6) / 2


def hb_h4_ZZ():
    """The ZZ-geometry honeycomb plaquette, sites 1-2-3-4 around it (honeycomb.jl:46-68)."""
    II, Iy, yI = tout(sI, sI), tout(sI, sigmay), tout(sigmay, sI)
    Ix, xI, zz = tout(sI, sigmax), tout(sigmax, sI), tout(sigmaz, sigmaz)
    ts = tout_site
    H = -ts(ts(II, Ix), ts(xI, II)) - ts(ts(Ix, II), ts(II, xI)) - ts(ts(Iy, yI), ts(II, II)) \
        - ts(ts(II, II), ts(yI, Iy)) - ts(ts(II, zz), ts(II, II)) - ts(ts(II, II), ts(II, zz))
    d = II.shape[0]
    return H.reshape(d * d, d * d, d * d, d * d, order="F") / 2


def _svd_sqrt(x):
    # wrap_svd (src/utility/wrapper.jl:25-55, cutoff 1e-12, no cap), then u sqrt(s), sqrt(s) v^H
    u, s, vh = np.linalg.svd(x, full_matrices=False)
    n = int((s / s.max() > 1e-12).sum())
    if s.size > n and abs(s[n] - s[n - 1]) < 1e-10:
        n += 1
    r = np.sqrt(s[:n])
    return u[:, :n] * r, r[:, None] * vh[:n]


def get_local_h(h):
    """local_h.jl:1-31: [s1 (d,d,r1), s2 (r1,d,d,r2), s3 (r2,d,d,r3),
    s4 (r3,d,d)] with h[(k1 b1), (k2 b2), (k3 b3), (k4 b4)] = s1 s2 s3 s4."""
    h = np.asarray(h)
    dims = h.shape
    d = int(round(math.sqrt(dims[0])))
    sl = []
    s1, sd = _svd_sqrt(h.reshape(dims[0], -1, order="F"))
    sl.append(s1.reshape(d, d, -1, order="F"))
    for i in range(len(dims) - 2):
        dl = sd.shape[0]
        si, sd = _svd_sqrt(sd.reshape(dl * dims[i], -1, order="F"))
        sl.append(si.reshape(dl, d, d, -1, order="F"))
    dl = sd.shape[0]
    sl.append(sd.reshape(dl, d, d, order="F"))
    return sl
