"""Learned spin-up proposal (no counterpart in the reference).

A developed turbulent (u, v, w, c) state at a given wave phase is proposed by
a factorized spectral neural operator (runtime.SpectralOperatorNet: BF16 MFMA
forward/backward on the device), projected divergence-free, and handed to the
plain DNS, which then runs unchanged.  The network only replaces the
initial condition; every time step is the DNS's own.

Representation: the network works on a reduced grid -- nx_r x ny_r Fourier
points (the modes |kx|, |ky| < nx_r/2 of the DNS) times nz_r Chebyshev
Gauss-Lobatto points (the first nz_r Chebyshev coefficients of each
wall-normal profile).  Fields are [z][y][x][field], bottom first.

Training pairs: target = a DNS snapshot at phase phi minus the laminar
Stokes state at phi; input = the snapshot's own deviation from its phase
mean scaled to the amplitude of the plain start's noise, plus the phase
(cos phi, sin phi) and z.  At inference the input is the plain start minus
the laminar state (its noise), and the proposal is the laminar state plus
the network's output.
"""
import json
import math
import os

import numpy as np

from . import runtime as RT
from . import workloads as W

GRID = dict(nx=24, ny=24, nz=33)
NFIELD = 4            # u, v, w, c
CIN = NFIELD + 3      # + cos(phase), sin(phase), z


# ---------------------------------------------------------------------------
# Chebyshev transforms on Gauss-Lobatto points (bottom first)
# ---------------------------------------------------------------------------
def cheb_analysis(n):
    """Matrix taking values at the n CGL points to Chebyshev coefficients."""
    N = n - 1
    j = np.arange(n)
    z = -np.cos(math.pi * j / N)
    k = np.arange(n)
    T = np.cos(np.outer(k, np.arccos(np.clip(z, -1.0, 1.0))))
    cj = np.where((j == 0) | (j == N), 2.0, 1.0)
    ck = np.where((k == 0) | (k == N), 2.0, 1.0)
    return 2.0 / N * T / np.outer(ck, cj)


def cheb_synthesis(ncoef, z):
    """Matrix evaluating the first ncoef Chebyshev polynomials at points z."""
    return np.cos(np.outer(np.arccos(np.clip(z, -1.0, 1.0)), np.arange(ncoef)))


def _ky_index(nyc, kmax):
    """DNS spectral y index of ky = -kmax .. kmax (FFT order in the DNS)."""
    return [ky if ky >= 0 else nyc + ky for ky in range(-kmax, kmax + 1)]


# ---------------------------------------------------------------------------
# DNS state <-> network grid
# ---------------------------------------------------------------------------
def reduce_state(uvw, c, grid=GRID):
    """DNS spectral fields uvw (3, mx, nyc, nz), c (1, mx, nyc, nz) ->
    values on the reduced grid, shape (nz_r, ny_r, nx_r, 4)."""
    spec = np.concatenate([uvw, c], axis=0)
    _, mx, nyc, nz = spec.shape
    nxr, nyr, nzr = grid["nx"], grid["ny"], grid["nz"]
    kmax = min(nxr, nyr) // 2 - 1
    assert mx > kmax and nyc // 2 > kmax, "DNS grid too coarse for the reduced grid"
    sub = spec[:, :kmax + 1, _ky_index(nyc, kmax), :]              # (4, kx, ky, nz)
    # wall-normal: DNS values -> coefficients -> the first nz_r at the reduced points
    A = cheb_analysis(nz)[:nzr]                                     # (nzr, nz)
    zr = -np.cos(math.pi * np.arange(nzr) / (nzr - 1))
    Sy = cheb_synthesis(nzr, zr)                                    # (nzr, nzr)
    coef = np.einsum("kj,fxyj->fxyk", A, sub)
    vals = np.einsum("ik,fxyk->fxyi", Sy, coef)                     # (4, kx, ky, nzr)
    G = np.zeros((NFIELD, nzr, nyr, nxr // 2 + 1), dtype=np.complex128)
    for j, ky in enumerate(range(-kmax, kmax + 1)):
        G[:, :, ky % nyr, :kmax + 1] = np.transpose(vals[:, :, j, :], (0, 2, 1))
    phys = np.fft.irfft2(G, s=(nyr, nxr), axes=(-2, -1)) * (nxr * nyr)
    return np.transpose(phys, (1, 2, 3, 0)).copy()                 # (z, y, x, f)


def expand_state(vals, mx, nyc, nz, grid=GRID):
    """Reduced-grid values (nz_r, ny_r, nx_r, 4) -> DNS spectral uvw, c with
    the modes |kx|, |ky| <= nx_r/2 - 1 and nz_r Chebyshev coefficients."""
    nxr, nyr, nzr = grid["nx"], grid["ny"], grid["nz"]
    kmax = min(nxr, nyr) // 2 - 1
    F = np.fft.rfft2(np.transpose(vals, (3, 0, 1, 2)), axes=(-2, -1)) / (nxr * nyr)   # (4, nzr, nyr, nxr/2+1)
    A = cheb_analysis(nzr)
    z = W.cgl_grid(nz)
    Sy = cheb_synthesis(nzr, z)                                     # (nz, nzr)
    M = Sy @ A                                                      # reduced values -> DNS points
    spec = np.zeros((NFIELD, mx, nyc, nz), dtype=np.complex128)
    iy = _ky_index(nyc, kmax)
    for j, ky in enumerate(range(-kmax, kmax + 1)):
        prof = F[:, :, ky % nyr, :kmax + 1]                          # (4, nzr, kx)
        spec[:, :kmax + 1, iy[j], :] = np.einsum("jk,fkx->fxj", M, prof)
    # exact Hermitian symmetry on kx = 0 and real mean
    for j in range(1, kmax + 1):
        a = spec[:, 0, j, :]
        b = spec[:, 0, nyc - j, :]
        h = 0.5 * (a + np.conj(b))
        spec[:, 0, j, :] = h
        spec[:, 0, nyc - j, :] = np.conj(h)
    spec[:, 0, 0, :] = spec[:, 0, 0, :].real
    return spec[:3].copy(), spec[3:4].copy()


def laminar_grid(cfg, phase, grid=GRID):
    """Laminar Stokes state on the reduced grid at wave phase `phase` (u along
    the forcing direction, c = 1)."""
    nxr, nyr, nzr = grid["nx"], grid["ny"], grid["nz"]
    zr = -np.cos(math.pi * np.arange(nzr) / (nzr - 1))
    omega = cfg["force_amplitude"][1]
    top = "freeslip" if cfg["vt_u1"][3] == 1.0 else "noslip"
    prof = W.stokes_profile(cfg["reynolds"], zr, phase / omega, omega=omega, top=top)
    v = np.zeros((nzr, nyr, nxr, NFIELD))
    comp = 0 if cfg["force_direction"][1] == 1 else 1
    v[:, :, :, comp] = prof[:, None, None]
    v[:, :, :, 3] = 1.0
    return v


def network_input(vals, phase, grid=GRID):
    """(points, CIN) input: fields, cos/sin of the phase, z."""
    nzr = grid["nz"]
    zr = -np.cos(math.pi * np.arange(nzr) / (nzr - 1))
    x = np.zeros(vals.shape[:3] + (CIN,))
    x[..., :NFIELD] = vals
    x[..., NFIELD] = math.cos(phase)
    x[..., NFIELD + 1] = math.sin(phase)
    x[..., NFIELD + 2] = zr[:, None, None]
    return x.reshape(-1, CIN)


# ---------------------------------------------------------------------------
# model
# ---------------------------------------------------------------------------
class Proposal:
    """Network + normalisation.  Inputs and targets are normalised per field
    (mean, std over the training targets)."""

    def __init__(self, cfg, channels=32, layers=4, batch=8, grid=GRID, seed=0):
        self.cfg = cfg
        self.grid = dict(grid)
        self.net = RT.SpectralOperatorNet(nx=grid["nx"], ny=grid["ny"], nz=grid["nz"], channels=channels,
                                          layers=layers, cin=CIN, cout=NFIELD, batch=batch)
        self.batch = batch
        self.channels, self.layers = channels, layers
        self.mean = np.zeros(NFIELD)
        self.std = np.ones(NFIELD)
        self.noise_rms = 0.1
        # developed-flow TKE by wave phase (12 bins from phase 0); when set,
        # a proposal's fluctuations are scaled to it (see propose)
        self.tke_phase = None
        self.tke_lo_phase = None    # the part of it inside the reduced band
        self.net.set_params(init_params(self.net.nparams, grid, channels, layers, seed))

    def _norm_in(self, x):
        x = x.copy()
        x[:, :NFIELD] = (x[:, :NFIELD] - self.mean) / self.std
        return x

    # Inputs and targets are deviations from the laminar state at the phase,
    # normalised per field: the network learns the mean-profile correction and
    # the fluctuations, not the (known) laminar part.
    def training_pair(self, snap, phase, phase_mean):
        """(input, target) for one snapshot (reduced grid) at `phase`."""
        lam = laminar_grid(self.cfg, phase, self.grid)
        dev = snap - phase_mean
        rms = math.sqrt(float(np.mean(dev[..., :3] ** 2))) + 1e-30
        x = network_input(dev * (self.noise_rms / rms), phase, self.grid)
        t = ((snap - lam).reshape(-1, NFIELD) - self.mean) / self.std
        return self._norm_in(x), t

    def fit(self, snaps, phases, epochs=30, lr=2e-3, log=None, seed=0):
        """snaps (n, nz_r, ny_r, nx_r, 4), phases (n,) in radians."""
        snaps = np.asarray(snaps, dtype=np.float64)
        phases = np.asarray(phases, dtype=np.float64)
        devs = np.stack([s - laminar_grid(self.cfg, ph, self.grid) for s, ph in zip(snaps, phases)])
        self.mean = devs.reshape(-1, NFIELD).mean(0)
        self.std = devs.reshape(-1, NFIELD).std(0) + 1e-12
        # phase means (12 bins) of the snapshots
        bins = np.floor((np.mod(phases, 2 * math.pi)) / (2 * math.pi) * 12).astype(int) % 12
        pmean = np.stack([snaps[bins == b].mean(0) if np.any(bins == b) else snaps.mean(0) for b in range(12)])
        X = []
        T = []
        for s, ph, b in zip(snaps, phases, bins):
            x, t = self.training_pair(s, ph, pmean[b])
            X.append(x)
            T.append(t)
        X = np.asarray(X, dtype=np.float32)
        T = np.asarray(T, dtype=np.float32)
        w = np.ones(NFIELD, dtype=np.float32)
        rng = np.random.default_rng(seed)
        n = len(X)
        hist = []
        for ep in range(epochs):
            order = rng.permutation(n)
            tot = 0.0
            nb = 0
            lr_ep = lr * 0.5 * (1 + math.cos(math.pi * ep / max(1, epochs)))
            for i in range(0, n - self.batch + 1, self.batch):
                idx = order[i:i + self.batch]
                tot += self.net.train_step(X[idx], T[idx], w, lr_ep)
                nb += 1
            hist.append(tot / max(1, nb))
            if log:
                log("surrogate: epoch %d loss %.5f" % (ep, hist[-1]))
        return hist

    def propose_grid(self, vals_in, phase):
        lam = laminar_grid(self.cfg, phase, self.grid)
        # the start's deviation from the laminar state, scaled to the size the
        # training inputs had (any noise amplitude of the plain start)
        dev = vals_in - lam
        rms = math.sqrt(float(np.mean(dev[..., :3] ** 2)))
        if rms > 0:
            dev = dev * (self.noise_rms / rms)
        x = self._norm_in(network_input(dev, phase, self.grid)).astype(np.float32)[None]
        y = self.net.forward(x)[0]
        return lam + (y * self.std + self.mean).reshape(vals_in.shape)

    def propose(self, uvw, c, phase):
        """DNS spectral plain start -> proposed DNS spectral state (not yet
        projected; Solver.project() after upload makes it divergence-free)."""
        _, mx, nyc, nz = uvw.shape
        vals = reduce_state(uvw, c, self.grid)
        out = self.propose_grid(vals, phase)
        # wall values of the start: no-slip bottom, no flow through the top
        out[0, :, :, :3] = 0.0
        out[-1, :, :, 2] = 0.0
        # the proposal holds the reduced modes only: the start's noise above
        # them is not part of a developed state
        uvw_p, c_p = expand_state(out, mx, nyc, nz, self.grid)
        if self.tke_phase is not None:
            # A least-squares fit predicts the conditional mean of the
            # fluctuations, with less energy than any one realisation, and
            # the reduced grid holds none of the energy above its band.  The
            # band's fluctuations (every mode but the mean) are scaled to the
            # developed flow's band energy at this phase, and the modes above
            # the band are the start's own, scaled to the developed flow's
            # energy there (random phases, the start's wall window).
            k = int(round(phase / (2 * math.pi / len(self.tke_phase)))) % len(self.tke_phase)
            lo = self.tke_lo_phase[k] if self.tke_lo_phase is not None else self.tke_phase[k]
            e_lo = fluctuation_energy(uvw_p)
            if e_lo > 0:
                mean = uvw_p[:, 0, 0, :].copy()
                uvw_p *= math.sqrt(lo / e_lo)
                uvw_p[:, 0, 0, :] = mean
            if self.tke_lo_phase is not None:
                hi = self.tke_phase[k] - lo
                out_band = ~band_mask(mx, nyc, self.grid)
                small = np.where(out_band[None, :, :, None], uvw, 0)
                e_hi = fluctuation_energy(small)
                if e_hi > 0 and hi > 0:
                    uvw_p = uvw_p + small * math.sqrt(hi / e_hi)
        return uvw_p, c_p

    def save(self, path, record=None):
        """Parameters and normalisation, plus `record` (JSON-able: how the
        model was made) in the metadata."""
        np.savez(path, params=self.net.params(), mean=self.mean, std=self.std,
                 meta=json.dumps(dict(grid=self.grid, channels=self.channels, layers=self.layers,
                                      noise_rms=self.noise_rms, tke_phase=self.tke_phase,
                                      tke_lo_phase=self.tke_lo_phase, record=record)))

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        self.net.set_params(d["params"])
        self.mean = d["mean"]
        self.std = d["std"]
        meta = json.loads(str(d["meta"]))
        self.noise_rms = meta.get("noise_rms", self.noise_rms)
        self.tke_phase = meta.get("tke_phase")
        self.tke_lo_phase = meta.get("tke_lo_phase")


def init_params(n, grid, C, L, seed):
    """Fan-in scaled random parameters in the device layout (lift, layers,
    projection; the per-mode mixing starts small)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n, dtype=np.float32)
    mxm, mym, nz = grid["nx"] // 2 + 1, grid["ny"] // 2 + 1, grid["nz"]
    o = 0

    def put(size, scale):
        nonlocal o
        p[o:o + size] = rng.standard_normal(size) * scale
        o += size

    put(CIN * C, 1.0 / math.sqrt(CIN))
    o += C
    for _ in range(L):
        put(C * C, 0.5 / math.sqrt(C))
        o += C
        put(mxm * 2 * C * C, 0.25 / math.sqrt(C))
        put(mym * 2 * C * C, 0.25 / math.sqrt(C))
        put(nz * C * C, 0.25 / math.sqrt(C))
    put(C * NFIELD, 1.0 / math.sqrt(C))
    o += NFIELD
    assert o == n
    return p


# ---------------------------------------------------------------------------
# statistics used to assess a start
# ---------------------------------------------------------------------------
def fluctuation_energy(uvw):
    """Turbulent kinetic energy of the non-mean modes of a spectral velocity
    (Parseval on the retained modes, kx > 0 counted twice), averaged over the
    wall-normal points."""
    fl = uvw.copy()
    fl[:, 0, 0, :] = 0
    wgt = np.ones(uvw.shape[1])
    wgt[1:] = 2.0
    return 0.5 * float(np.sum(wgt[None, :, None, None] * np.abs(fl) ** 2)) / uvw.shape[-1]


def band_mask(mx, nyc, grid=GRID):
    """(mx, nyc) mask of the DNS modes the reduced grid holds (|kx|, |ky| <=
    kmax)."""
    kmax = min(grid["nx"], grid["ny"]) // 2 - 1
    m = np.zeros((mx, nyc), dtype=bool)
    m[:kmax + 1, _ky_index(nyc, kmax)] = True
    return m


def band_energy(uvw, grid=GRID):
    """fluctuation_energy of the modes inside the reduced band."""
    m = band_mask(uvw.shape[1], uvw.shape[2], grid)
    return fluctuation_energy(np.where(m[None, :, :, None], uvw, 0))


def flow_stats(uvw, c, feat):
    """Wall shear of the mean mode (dU/dz, dV/dz at the bottom), turbulent
    kinetic energy of the non-mean modes, mean concentration."""
    return dict(tau=float(math.hypot(feat[10], feat[11])), tke=fluctuation_energy(uvw),
                tke_lo=band_energy(uvw), cmean=float(np.real(c[0, 0, 0, :]).mean()))
