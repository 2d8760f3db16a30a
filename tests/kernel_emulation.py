"""CPU emulation of the HIP step's algebra (test infrastructure).

Same sequence as the kernels -- spectral-resident state, vorticity from D1 contractions,
c2r/product/r2c with mask, eigenbasis Helmholtz and Poisson with the library's own operator
matrices (dns_debug_operator, host-only), all wall-normal contractions in the kernels' even/odd
folded form (x_k +- x_{n-1-k} panels, two half-size blocks, outputs unfolded), mode (0,0) by
M0 -- in numpy.  Agreement with the oracle pins the algebra and the folded operators;
disagreement between this and the GPU is an implementation bug.
"""
import ctypes

import numpy as np

from oracle import oracle as orc


def operator(nz, dt, re, cn, exact, which):
    import f90_dns_amd as dns
    L = dns.lib()
    L.dns_debug_operator.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]
    r, c = ctypes.c_int(), ctypes.c_int()
    assert L.dns_debug_operator(nz, dt, re, cn, exact, which, None, ctypes.byref(r), ctypes.byref(c)) == 0
    out = np.empty(r.value * c.value)
    assert L.dns_debug_operator(nz, dt, re, cn, exact, which, ctypes.c_void_p(out.ctypes.data), None, None) == 0
    return out.reshape(r.value, c.value) if c.value > 1 else out


def fold(x):
    """(X_E, X_O) of a column stack x (axis 0 of length n): x_c +- x_{n-1-c}, centre in X_E."""
    n = x.shape[0]
    h = n // 2
    top, bot = x[:h], x[::-1][:h]
    xe = top + bot
    if n % 2:
        xe = np.concatenate([xe, x[h:h + 1]], axis=0)
    return xe, top - bot


def pad(x, rows):
    if x.shape[0] == rows:
        return x
    z = np.zeros((rows - x.shape[0],) + x.shape[1:], dtype=x.dtype)
    return np.concatenate([x, z], axis=0)


def unfold(g0, g1, n):
    """out_i = g0_i + g1_i, out_{n-1-i} = g0_i - g1_i (i < n/2), out_centre = g0_centre."""
    h = n // 2
    out = np.empty((n,) + g0.shape[1:], dtype=np.result_type(g0, g1))
    out[:h] = g0[:h] + g1[:h]
    out[n - 1 - np.arange(h)] = g0[:h] - g1[:h]
    if n % 2:
        out[h] = g0[h]
    return out


def mm(M, X):
    return np.einsum("jk,kyx->jyx", M, X)


class Emulator:
    def __init__(self, **kw):
        p = orc.make_params(**kw)
        self.p = p
        nx, ny, nz = p.nx, p.ny, p.nz
        self.nx, self.ny, self.nz = nx, ny, nz
        cn, ex = p.use_crank_nicolson, p.lgl_derivative
        op = lambda w: operator(nz, p.dt, p.re, cn, ex, w)
        self.hF, self.hB, self.hc1, self.hc2, self.hl = op(0), op(1), op(2), op(3), op(4)
        self.pF, self.pB, self.pDB, self.pl, self.M0 = op(5), op(6), op(7), op(8), op(9)
        self.D1f, self.wz = op(10), op(11)
        self.D1 = orc.diff_matrix_exact(nz) if ex else orc.diff_matrix(nz)
        self.f = (0.5 if cn else 1.0) * p.dt / p.re
        xw, yw = orc.wavenumbers(nx, ny, p.xlen, p.ylen)
        self.kx = xw[None, None, :]
        self.ky = yw[None, :, None]
        self.k2 = self.kx ** 2 + self.ky ** 2
        kxe = xw.copy(); kxe[nx // 2] = 0.0
        kye = np.broadcast_to(yw[:, None], (ny, nx // 2 + 1)).copy()
        kye[ny // 2, 0] = 0.0; kye[ny // 2, nx // 2] = 0.0
        self.kxe = kxe[None, None, :]
        self.kye = kye[None, :, :]
        self.mask = orc.dealias_mask(nx, ny).astype(bool)[None]
        self.G = p.target_pressure_gradient
        self.err_int = self.err_prev = self.gsum = 0.0
        self.gcount = 0

    def set_fields(self, u, v, w, pr):
        self.uh, self.vh, self.wh, self.ph = (orc.fft_r2c(a) for a in (u, v, w, pr))

    def get_fields(self):
        return [orc.fft_c2r(a, self.nx) for a in (self.uh, self.vh, self.wh, self.ph)]

    def dz(self, a):
        """D1 a as the folded contraction: s = D1 x_odd on X_O, r = D1 x_even on X_E."""
        nz = self.nz
        LE = nz - nz // 2
        xe, xo = fold(a)
        s = mm(self.D1f[:LE], pad(xo, LE))
        r = mm(self.D1f[LE:], xe)
        return unfold(s, r, nz)

    def step(self):
        p, nx, ny, nz, f, dt = self.p, self.nx, self.ny, self.nz, self.f, self.p.dt
        uh, vh, wh = self.uh, self.vh, self.wh
        ox = 1j * self.ky * wh - self.dz(vh)
        oy = self.dz(uh) - 1j * self.kx * wh
        oz = 1j * self.kx * vh - 1j * self.ky * uh
        u, v, w, wx, wy, wz_ = (orc.fft_c2r(a, nx) for a in (uh, vh, wh, ox, oy, oz))
        S = [v * wz_ - w * wy, w * wx - u * wz_, u * wy - v * wx]
        Sh = [orc.fft_r2c(s) for s in S]
        if p.use_dealias_23_rule:
            Sh = [np.where(self.mask, s, 0) for s in Sh]
        if p.flow_control_method == 2 and p.istart % p.controller_update_freq == 0:
            Ub = np.dot(self.wz, uh[:, 0, 0].real) / (nx * ny * self.wz.sum())
            e = p.target_bulk_velocity - Ub
            gu = 3.0 / p.re + p.controller_gain * e + 0.05 * p.controller_gain * self.err_int
            gc = max(0.0, min(10.0, gu))
            if abs(gc - gu) < 1e-12:
                self.err_int += e * dt if p.istart <= p.controller_update_freq else 0.5 * dt * (self.err_prev + e)
            self.err_prev = e
            self.G = gc
        elif p.flow_control_method == 1:
            self.G = p.target_pressure_gradient
        Sh[0] = Sh[0].copy()
        Sh[0][:, 0, 0] += self.G * nx * ny
        m = nz - 2
        star = []
        for c, (ah, sh) in enumerate(zip((uh, vh, wh), Sh)):
            ke = self.kxe if c == 0 else self.kye
            if c < 2:
                bc1 = -dt * (1j * ke * self.ph[0:1])
                bc2 = -dt * (1j * ke * self.ph[nz - 1:nz])
            else:
                bc1 = bc2 = np.zeros_like(ah[0:1])
            alpha = (1.0 - f * self.k2) if p.use_crank_nicolson else 1.0
            X1 = alpha * ah[1:nz - 1] + dt * sh[1:nz - 1]
            nE, LE = m - m // 2, nz - nz // 2
            ve, vo = fold(X1)
            ue, uo = fold(ah)
            P0 = np.concatenate([ve, ue], axis=0)
            P1 = np.concatenate([pad(vo, nE), pad(uo, LE)], axis=0)
            y = np.concatenate([mm(self.hF[:nE], P0), mm(self.hF[nE:], P1)], axis=0)
            y = y + self.hc1[:, None, None] * bc1 + self.hc2[:, None, None] * bc2
            z = y / (self.hl[:, None, None] + f * self.k2)
            out = np.empty_like(ah)
            pp = mm(self.hB[:nE], z[:nE])
            qq = mm(self.hB[nE:], pad(z[nE:], nE))
            out[1:nz - 1] = unfold(pp, qq, m)
            out[0] = bc1[0]; out[nz - 1] = bc2[0]
            star.append(out)
        us, vs, ws = star
        div = 1j * self.kxe * us + 1j * self.kye * vs
        LE = nz - nz // 2
        de, do = fold(div)
        we, wo = fold(ws)
        P0 = np.concatenate([de, pad(wo, LE)], axis=0)
        P1 = np.concatenate([pad(do, LE), we], axis=0)
        y = np.concatenate([mm(self.pF[:LE], P0), mm(self.pF[LE:], P1)], axis=0)
        k2 = self.k2.copy(); k2[0, 0, 0] = 1.0
        z = -y / (dt * (self.pl[:, None, None] + k2))
        zE, zO = z[:LE], pad(z[LE:], LE)
        phi = unfold(mm(self.pB[:LE], zE), mm(self.pB[LE:], zO), nz)
        dphi = unfold(mm(self.pDB[:LE], zO), mm(self.pDB[LE:], zE), nz)
        b = self.wz * self.dz(ws)[:, 0, 0] / dt
        phi0 = np.concatenate([[0.0], self.M0 @ b[1:]])
        phi[:, 0, 0] = phi0
        dphi[:, 0, 0] = self.D1 @ phi0
        self.uh = us - dt * 1j * self.kxe * phi
        self.vh = vs - dt * 1j * self.kye * phi
        self.wh = ws - dt * dphi
        self.ph = self.ph + phi
        for a in (self.uh, self.vh, self.wh):
            a[0] = 0; a[nz - 1] = 0
