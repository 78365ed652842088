"""CPU restatement (numpy, FP64) of the spin-up proposal network.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nothing else.  The network
has no counterpart in the reference; this module restates, with plain
einsum/matmul in FP64, the definition the device kernels implement:

    H_0     = X P + p                                   (lift)
    S_l     = H_l W_l + b_l + Kx_l(H_l) + Ky_l(H_l) + Kz_l(H_l)
    H_{l+1} = H_l + gelu(S_l)
    Y       = H_L Q + q
    loss    = mean over samples, points, outputs of w_o (Y - T)^2

with K_a(H) = Dinv_a Mix_a(D_a H) along the lines of axis a (x, y: real DFT,
modes 0..n/2 as interleaved real and imaginary rows; z: Chebyshev transform
on the Gauss-Lobatto points), Mix_a a per-mode channel mixing in row-vector
form g = f R (complex for x, y: R = A + iB; real for z).  Gradients are
written out by hand (reverse mode) and pinned in the tests by central finite
differences of this same forward.
"""
import math

import numpy as np


def modes(n):
    return n // 2 + 1


def dft_mats(n):
    mm = modes(n)
    D = np.zeros((2 * mm, n))
    Di = np.zeros((n, 2 * mm))
    x = np.arange(n)
    for m in range(mm):
        th = 2.0 * math.pi * m * x / n
        D[2 * m] = np.cos(th) / n
        D[2 * m + 1] = -np.sin(th) / n
        cm = 1.0 if (m == 0 or 2 * m == n) else 2.0
        Di[:, 2 * m] = cm * np.cos(th)
        Di[:, 2 * m + 1] = -cm * np.sin(th)
    return D, Di


def cheb_mats(n):
    N = n - 1
    j = np.arange(n)
    z = -np.cos(math.pi * j / N)
    k = np.arange(n)
    T = np.cos(np.outer(k, np.arccos(np.clip(z, -1, 1))))   # T[k, j]
    cj = np.where((j == 0) | (j == N), 2.0, 1.0)
    ck = np.where((k == 0) | (k == N), 2.0, 1.0)
    D = 2.0 / N * T / np.outer(ck, cj)
    Di = T.T.copy()
    return D, Di


class Layout:
    """Offsets of the parameter block (same order as the device network)."""

    def __init__(self, nx, ny, nz, C, L, cin, cout):
        self.nx, self.ny, self.nz, self.C, self.L, self.cin, self.cout = nx, ny, nz, C, L, cin, cout
        o = 0
        self.P = o; o += cin * C
        self.pb = o; o += C
        self.W, self.b, self.Rx, self.Ry, self.Rz = [], [], [], [], []
        for _ in range(L):
            self.W.append(o); o += C * C
            self.b.append(o); o += C
            self.Rx.append(o); o += modes(nx) * 2 * C * C
            self.Ry.append(o); o += modes(ny) * 2 * C * C
            self.Rz.append(o); o += nz * C * C
        self.Q = o; o += C * cout
        self.qb = o; o += cout
        self.total = o

    def views(self, p):
        C, cin, cout = self.C, self.cin, self.cout
        v = dict(P=p[self.P:self.P + cin * C].reshape(cin, C), pb=p[self.pb:self.pb + C],
                 Q=p[self.Q:self.Q + C * cout].reshape(C, cout), qb=p[self.qb:self.qb + cout])
        mx, my = modes(self.nx), modes(self.ny)
        v["W"] = [p[o:o + C * C].reshape(C, C) for o in self.W]
        v["b"] = [p[o:o + C] for o in self.b]
        v["Rx"] = [p[o:o + mx * 2 * C * C].reshape(mx, 2, C, C) for o in self.Rx]
        v["Ry"] = [p[o:o + my * 2 * C * C].reshape(my, 2, C, C) for o in self.Ry]
        v["Rz"] = [p[o:o + self.nz * C * C].reshape(self.nz, C, C) for o in self.Rz]
        return v


def init_params(lay, seed=0, scale=1.0):
    """Random initial parameters (fan-in scaled)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(lay.total)
    v = lay.views(p)
    C = lay.C
    v["P"][:] = rng.standard_normal(v["P"].shape) * scale / math.sqrt(lay.cin)
    v["Q"][:] = rng.standard_normal(v["Q"].shape) * scale / math.sqrt(C)
    for l in range(lay.L):
        v["W"][l][:] = rng.standard_normal((C, C)) * scale * 0.5 / math.sqrt(C)
        v["Rx"][l][:] = rng.standard_normal(v["Rx"][l].shape) * scale * 0.25 / math.sqrt(C)
        v["Ry"][l][:] = rng.standard_normal(v["Ry"][l].shape) * scale * 0.25 / math.sqrt(C)
        v["Rz"][l][:] = rng.standard_normal(v["Rz"][l].shape) * scale * 0.25 / math.sqrt(C)
    return p


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(s):
    return 0.5 * s * (1.0 + _erf(s / math.sqrt(2.0)))


def gelu_d(s):
    return 0.5 * (1.0 + _erf(s / math.sqrt(2.0))) + s * np.exp(-0.5 * s * s) / math.sqrt(2.0 * math.pi)


class Net:
    """H[b, z, y, x, c] activations, FP64."""

    def __init__(self, nx=24, ny=24, nz=33, C=32, L=4, cin=7, cout=4):
        self.lay = Layout(nx, ny, nz, C, L, cin, cout)
        self.Dx, self.Dxi = dft_mats(nx)
        self.Dy, self.Dyi = dft_mats(ny)
        self.Dz, self.Dzi = cheb_mats(nz)

    # --- axis operator: S += Dinv Mix(D H) -------------------------------
    def _mats(self, axis):
        return [(self.Dx, self.Dxi), (self.Dy, self.Dyi), (self.Dz, self.Dzi)][axis]

    @staticmethod
    def _to_axis(H, axis):
        # move the axis (x: 3, y: 2, z: 1) next to channels: [..., n, c]
        src = {0: 3, 1: 2, 2: 1}[axis]
        return np.moveaxis(H, src, -2)

    @staticmethod
    def _from_axis(A, axis):
        src = {0: 3, 1: 2, 2: 1}[axis]
        return np.moveaxis(A, -2, src)

    def _mix(self, F, R, axis):
        # F [..., rows, C]; complex axes: rows 2m (re), 2m+1 (im)
        if axis == 2:
            return np.einsum("...kc,kcd->...kd", F, R)
        fre, fim = F[..., 0::2, :], F[..., 1::2, :]
        A, B = R[:, 0], R[:, 1]
        gre = np.einsum("...mc,mcd->...md", fre, A) - np.einsum("...mc,mcd->...md", fim, B)
        gim = np.einsum("...mc,mcd->...md", fre, B) + np.einsum("...mc,mcd->...md", fim, A)
        G = np.empty_like(F)
        G[..., 0::2, :] = gre
        G[..., 1::2, :] = gim
        return G

    def _mix_back(self, F, dG, R, axis):
        """dF and dR of G = Mix(F, R)."""
        if axis == 2:
            dF = np.einsum("...kd,kcd->...kc", dG, R)
            dR = np.einsum("lkc,lkd->kcd", F.reshape((-1,) + F.shape[-2:]), dG.reshape((-1,) + dG.shape[-2:]))
            return dF, dR
        fre, fim = F[..., 0::2, :], F[..., 1::2, :]
        gre, gim = dG[..., 0::2, :], dG[..., 1::2, :]
        A, B = R[:, 0], R[:, 1]
        dfre = np.einsum("...md,mcd->...mc", gre, A) + np.einsum("...md,mcd->...mc", gim, B)
        dfim = -np.einsum("...md,mcd->...mc", gre, B) + np.einsum("...md,mcd->...mc", gim, A)
        fl = lambda a: a.reshape((-1,) + a.shape[-2:])
        fre, fim, gre, gim = fl(fre), fl(fim), fl(gre), fl(gim)
        dA = np.einsum("lmc,lmd->mcd", fre, gre) + np.einsum("lmc,lmd->mcd", fim, gim)
        dB = -np.einsum("lmc,lmd->mcd", fim, gre) + np.einsum("lmc,lmd->mcd", fre, gim)
        dF = np.empty_like(F)
        dF[..., 0::2, :] = dfre
        dF[..., 1::2, :] = dfim
        return dF, np.stack([dA, dB], axis=1)

    def _axis(self, H, R, axis):
        D, Di = self._mats(axis)
        Ha = self._to_axis(H, axis)
        F = np.einsum("kn,...nc->...kc", D, Ha)
        G = self._mix(F, R, axis)
        return self._from_axis(np.einsum("nk,...kc->...nc", Di, G), axis)

    def _axis_back(self, H, R, dS, axis):
        D, Di = self._mats(axis)
        Ha = self._to_axis(H, axis)
        F = np.einsum("kn,...nc->...kc", D, Ha)
        dG = np.einsum("nk,...nc->...kc", Di, self._to_axis(dS, axis))
        dF, dR = self._mix_back(F, dG, R, axis)
        dH = self._from_axis(np.einsum("kn,...kc->...nc", D, dF), axis)
        return dH, dR

    # --- network ------------------------------------------------------------
    def shape5(self, nb, c):
        return (nb, self.lay.nz, self.lay.ny, self.lay.nx, c)

    def forward(self, p, x, keep=False):
        """x [nb, points, cin] -> y [nb, points, cout]."""
        lay = self.lay
        v = lay.views(np.asarray(p, dtype=np.float64))
        nb = x.shape[0]
        X = np.asarray(x, dtype=np.float64).reshape(self.shape5(nb, lay.cin))
        H = X @ v["P"] + v["pb"]
        Hs, Ss = [], []
        for l in range(lay.L):
            S = H @ v["W"][l] + v["b"][l]
            S = S + self._axis(H, v["Rx"][l], 0) + self._axis(H, v["Ry"][l], 1) + self._axis(H, v["Rz"][l], 2)
            Hs.append(H)
            Ss.append(S)
            H = H + gelu(S)
        Y = H @ v["Q"] + v["qb"]
        y = Y.reshape(nb, -1, lay.cout)
        if keep:
            return y, (X, Hs, Ss, H)
        return y

    def loss(self, p, x, t, w):
        y = self.forward(p, x)
        return float(np.mean(np.asarray(w)[None, None, :] * (y - t) ** 2))

    def loss_grad(self, p, x, t, w):
        lay = self.lay
        v = lay.views(np.asarray(p, dtype=np.float64))
        y, (X, Hs, Ss, HL) = self.forward(p, x, keep=True)
        nb = x.shape[0]
        w = np.asarray(w, dtype=np.float64)
        e = y - t
        count = e.size
        loss = float(np.sum(w[None, None, :] * e * e) / count)
        dY = (2.0 * w[None, None, :] * e / count).reshape(self.shape5(nb, lay.cout))
        g = np.zeros(lay.total)
        gv = lay.views(g)
        gv["Q"][:] = HL.reshape(-1, lay.C).T @ dY.reshape(-1, lay.cout)
        gv["qb"][:] = dY.reshape(-1, lay.cout).sum(0)
        dH = dY @ v["Q"].T
        for l in reversed(range(lay.L)):
            H, S = Hs[l], Ss[l]
            dS = dH * gelu_d(S)
            gv["W"][l][:] = H.reshape(-1, lay.C).T @ dS.reshape(-1, lay.C)
            gv["b"][l][:] = dS.reshape(-1, lay.C).sum(0)
            dHn = dH + dS @ v["W"][l].T
            for ax, key in ((0, "Rx"), (1, "Ry"), (2, "Rz")):
                dHa, dR = self._axis_back(H, v[key][l], dS, ax)
                dHn = dHn + dHa
                gv[key][l][:] = dR
            dH = dHn
        gv["P"][:] = X.reshape(-1, lay.cin).T @ dH.reshape(-1, lay.C)
        gv["pb"][:] = dH.reshape(-1, lay.C).sum(0)
        return loss, g
