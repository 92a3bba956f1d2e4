"""Spectral Laplacians of the backprojection step -- a numpy restatement.

TEST INFRASTRUCTURE ONLY: the checker of the product's wrt_laplace2d / wrt_laplace3d.  Only
tests/ may use it.

laplace2dfft.m / laplace3dfft.m (radon_inversion/backprojection/laplace2dfft.m:1-23,
laplace3dfft.m:1-25): output = ifftn(U .* fftn(u)) with the signed integer frequencies
k (k <= ngrid/2 kept as k, else k - ngrid) and U = -(2 pi / period)^2 * sum of k^2 over the
axes.  The operator is the same along every axis, so the array's index order does not
matter.  The reference returns Octave's complex ifftn; its imaginary part is rounding noise
for real u, and the real part is returned here.
"""
import numpy as np


def frequencies(ngrid):
    """the signed frequency of each DFT index, as laplace3dfft.m:8-15 builds it"""
    k = np.arange(ngrid)
    return np.where(k <= ngrid / 2, k, k - ngrid).astype(np.float64)


def laplace_fft(u, period):
    """spectral Laplacian of a real 2-D or 3-D array with equal sides"""
    u = np.asarray(u, dtype=np.float64)
    n = u.shape[0]
    assert all(s == n for s in u.shape) and u.ndim in (2, 3)
    f = frequencies(n)
    grids = np.meshgrid(*([f] * u.ndim), indexing="ij")
    U = -(2 * np.pi / period) ** 2 * sum(g ** 2 for g in grids)
    return np.real(np.fft.ifftn(U * np.fft.fftn(u)))
