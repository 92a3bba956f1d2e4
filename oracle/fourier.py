"""Fourier-slice (Chang-type) inversion of Radon data -- a numpy restatement.

TEST INFRASTRUCTURE ONLY: the checker of the product's wrt_fourier_inversion.  Only tests/
may use it.

rtft3d.m (radon_inversion/rtft3d.m:3-137): for every direction (phi, theta) the Radon data
along s are zero-padded to ntotal = padding * floor(1/dshift) points, dshift = 2 rsupp /
(nshift - 1), Fourier transformed (fft), multiplied by exp(2 pi i rsupp f_k) with the
uncentred frequencies f_k = k dfreq, dfreq = 1 / (ntotal dshift), and fftshift-ed; the node of
value m is rho_m u(phi, theta), rho_m = (m - ntotal/2) dfreq (fftshift of
centered_frequencies.m), with the jacobian weight 0.5 dshift dfreq dphi w_theta rho_m^2.  The
appended Cartesian nodes outside the Nyquist ball carry the value 0 and add nothing below.

nfft_reconstruct_3d.m (:19-53): the adjoint NFFT of weight * value at the nodes scaled by
dx = 2/ngrid, i.e. f(p) = sum_j w_j v_j exp(2 pi i p . x_j) at p = k dx, k in
[-ngrid/2, ngrid/2)^3, real part.  Here the sum is evaluated exactly (the NFFT library
approximates it), direction by direction.  chang3d.m (:3-30) divides by w0.

In 2D (rtft2d.m:3-62, nfft_reconstruct_2d.m:19-45, chang2d.m:3-26) the data are
radon[s][phi], the directions (cos phi, sin phi), the values dshift * fftshift(fft .* corr)
and the weights 0.5 dfreq dphi |rho_m|; the output is out[ix][iy] = test_function(iy, ix).

Output order out[ix][iy][iz] at (x, y, z) = ((ix, iy, iz) - ngrid/2) * 2/ngrid; the
reference's two permutes (nfft_reconstruct_3d.m:48-49) give the same array as
test_function(iy, ix, iz) in Octave.
"""
import numpy as np


def rtft3d(radon, phi, theta, theta_w, rsupp=1.0, padding=4.0):
    """radon[s][phi][theta] -> (rho[ntotal], dirs[ndir][3], coef[ndir][ntotal]) with
    coef = jacobian weight * value; direction d = iphi * ntheta + itheta"""
    radon = np.asarray(radon, dtype=np.float64)
    nshift, nphi, ntheta = radon.shape
    dphi = 2 * np.pi / nphi
    dshift = 2 * rsupp / (nshift - 1)
    ntotal = int(round(padding * np.floor(1 / dshift)))
    dfreq = 1 / (ntotal * dshift)
    freq = np.arange(ntotal) * dfreq
    corr = np.exp(2 * np.pi * 1j * rsupp * freq)
    rho = (np.arange(ntotal) - ntotal // 2) * dfreq
    sig = radon.reshape(nshift, nphi * ntheta).T  # [dir][s]
    X = np.fft.fft(sig, ntotal, axis=1)
    V = np.fft.fftshift(X * corr[None, :], axes=1)
    w = np.tile(np.asarray(theta_w, dtype=np.float64), nphi)  # per direction
    jw = 0.5 * dshift * dfreq * dphi * w[:, None] * rho[None, :] ** 2
    P, T = np.meshgrid(phi, theta, indexing="ij")
    P, T = P.ravel(), T.ravel()
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=1)
    return rho, dirs, jw * V


def adjoint_sum(rho, dirs, coef, ngrid):
    """exact f(p) = Re sum_d sum_m coef[d, m] exp(2 pi i rho_m (p . u_d))"""
    x = (np.arange(ngrid) - ngrid // 2) * (2.0 / ngrid)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    out = np.zeros(pts.shape[0])
    for d in range(dirs.shape[0]):
        t = pts @ dirs[d]
        out += np.real(np.exp(2j * np.pi * np.outer(t, rho)) @ coef[d])
    return out.reshape(ngrid, ngrid, ngrid)


def chang3d(radon, phi, theta, theta_w, ngrid, w0=1.0, rsupp=1.0, padding=4.0):
    rho, dirs, coef = rtft3d(radon, phi, theta, theta_w, rsupp, padding)
    return adjoint_sum(rho, dirs, coef, ngrid) / w0


def rtft2d(radon, phi, rsupp=1.0, padding=4.0):
    """radon[s][phi] -> (rho[ntotal], dirs[nphi][2], coef[nphi][ntotal]), coef = weight * value"""
    radon = np.asarray(radon, dtype=np.float64)
    nshift, nphi = radon.shape
    dphi = 2 * np.pi / nphi
    dshift = 2 * rsupp / (nshift - 1)
    ntotal = int(round(padding * np.floor(1 / dshift)))
    dfreq = 1 / (ntotal * dshift)
    freq = np.arange(ntotal) * dfreq
    corr = np.exp(2 * np.pi * 1j * rsupp * freq)
    rho = (np.arange(ntotal) - ntotal // 2) * dfreq
    X = np.fft.fft(radon.T, ntotal, axis=1)
    V = dshift * np.fft.fftshift(X * corr[None, :], axes=1)
    jw = 0.5 * dfreq * dphi * np.abs(rho)
    phi = np.asarray(phi, dtype=np.float64)
    dirs = np.stack([np.cos(phi), np.sin(phi)], axis=1)
    return rho, dirs, jw[None, :] * V


def adjoint_sum2d(rho, dirs, coef, ngrid):
    """exact f(p) = Re sum_d sum_m coef[d, m] exp(2 pi i rho_m (p . u_d)), p on ngrid^2"""
    x = (np.arange(ngrid) - ngrid // 2) * (2.0 / ngrid)
    X, Y = np.meshgrid(x, x, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], axis=1)
    out = np.zeros(pts.shape[0])
    for d in range(dirs.shape[0]):
        t = pts @ dirs[d]
        out += np.real(np.exp(2j * np.pi * np.outer(t, rho)) @ coef[d])
    return out.reshape(ngrid, ngrid)


def chang2d(radon, phi, ngrid, w0=1.0, rsupp=1.0, padding=4.0):
    rho, dirs, coef = rtft2d(radon, phi, rsupp, padding)
    return adjoint_sum2d(rho, dirs, coef, ngrid) / w0
