"""Writes the golden vectors in this folder from numpy (pocketfft, LAPACK).

fft_golden.npz: plane transforms under the solver's truncation convention
  forward : numpy fft2 / (Nx*Ny), kept kx < mx, ky in FFT order (nyc modes)
  backward: zero padding, numpy ifft along y (unnormalised), numpy irfft along
            x (unnormalised; irfft discards the imaginary part of the DC bin,
            as FFTW's halfcomplex c2r does)
eig_golden.npz: eigenvalues (numpy.linalg.eigvals) of the interior
  Chebyshev operators with the boundary unknowns eliminated, for the
  boundary-row types the solver uses, built from the matrices the oracle
  exports are NOT used here: the operator is assembled from Trefethen's
  textbook CGL differentiation matrix (Spectral Methods in MATLAB, cheb.m).
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def trunc_forward(u):
    nv, nz, ny, nx = u.shape
    nxc, nyc = nx // 3 * 2, ny // 3 * 2
    mx = nxc // 2 + 1
    F = np.fft.fft2(u, axes=(-2, -1)) / (nx * ny)
    yf = [iy if iy < nyc // 2 else iy + ny - nyc for iy in range(nyc)]
    return F[:, :, yf, :mx].transpose(0, 3, 2, 1)


def trunc_backward(s, nx, ny):
    nv, mx, nyc, nz = s.shape
    yf = [iy if iy < nyc // 2 else iy + ny - nyc for iy in range(nyc)]
    full = np.zeros((nv, nz, ny, nx // 2 + 1), complex)
    full[:, :, yf, :mx] = s.transpose(0, 3, 2, 1)
    t = np.fft.ifft(full, axis=2) * ny
    return np.fft.irfft(t, n=nx, axis=3) * nx


def cheb_trefethen(N):
    """Trefethen's cheb(N): points x_j = cos(pi j/N) (decreasing), matrix D."""
    x = np.cos(np.pi * np.arange(N + 1) / N)
    c = np.hstack([2, np.ones(N - 1), 2]) * (-1) ** np.arange(N + 1)
    X = np.tile(x, (N + 1, 1)).T
    dX = X - X.T
    D = np.outer(c, 1 / c) / (dX + np.eye(N + 1))
    D = D - np.diag(D.sum(axis=1))
    return D, x


def interior_operator(nz, vt):
    """Bordered D2 with rows 0/N-1 = a f + b f' and the two boundary unknowns
    eliminated (grid ordered bottom z=-1 first)."""
    D, x = cheb_trefethen(nz - 1)
    P = np.eye(nz)[::-1]
    D1 = P @ D @ P            # increasing grid
    D2 = D1 @ D1
    C = D2.copy()
    C[0] = vt[1] * D1[0]
    C[-1] = vt[3] * D1[-1]
    C[0, 0] += vt[0]
    C[-1, -1] += vt[2]
    # eliminate y0, yN from the two boundary rows
    B = C[[0, -1]][:, [0, -1]]
    Binv = np.linalg.inv(B)
    inner = slice(1, nz - 1)
    M = C[inner, inner] - C[inner][:, [0, -1]] @ Binv @ C[[0, -1]][:, inner]
    return M


BC_TYPES = {
    "dirichlet_dirichlet": (1.0, 0.0, 1.0, 0.0),
    "dirichlet_neumann": (1.0, 0.0, 0.0, 1.0),
    "neumann_neumann": (0.0, 1.0, 0.0, 1.0),
    "robin_settling": (1e7 * 16200 * 1e-3, 1e7, 1e7 * 16200 * 1e-3, 1e7),
}


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for tag, shape in (("a", (2, 2, 32, 32)), ("b", (1, 3, 48, 96)), ("c", (1, 2, 6, 8))):
        u = rng.standard_normal(shape)
        out["fwd_in_" + tag] = u
        out["fwd_out_" + tag] = trunc_forward(u)
        nv, nz, ny, nx = shape
        nyc, mx = ny // 3 * 2, (nx // 3 * 2) // 2 + 1
        s = rng.standard_normal((nv, mx, nyc, nz)) + 1j * rng.standard_normal((nv, mx, nyc, nz))
        out["bwd_in_" + tag] = s
        out["bwd_out_" + tag] = trunc_backward(s, nx, ny)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)
    eig = {}
    for nz in (33, 65):
        for name, vt in BC_TYPES.items():
            M = interior_operator(nz, vt)
            ev = np.linalg.eigvals(M)
            eig["%s_%d" % (name, nz)] = np.sort_complex(ev)
    np.savez_compressed(os.path.join(HERE, "eig_golden.npz"), **eig)
    print("wrote fft_golden.npz, eig_golden.npz")


if __name__ == "__main__":
    main()
