"""Closed forms that pin the backprojection's orientation and output order and the
reduction's z stencil.  They are stated from the mathematics, not from the oracle's code, so
that the same checks run on the oracle (CPU) and through the HIP path (GPU).

Backprojection (radon_inversion/backprojection/C/src/integral.c:6-24, voxel loop
radon3d_ajoint.c:240-254).  Data g(s, phi, theta) = s * n_k(phi, theta), n the unit normal
(sin t cos p, sin t sin p, cos t): every row is linear in s, which a natural cubic spline
reproduces exactly, so the backprojection at a voxel centre x with |x| <= 1 (no cut, since
|x.n| <= |x|) is sum_phi sum_theta dphi w_theta (x.n) n_k = (4 pi / 3) x_k.  Uniform phi
(nphi >= 3) and Gauss-Legendre theta (ntheta >= 2) integrate the products n_j n_k exactly.
An axis permutation, a sign flip of a direction component or another output order moves
x_k to another axis or flips it.

Reduction (radon_reduction/slicing.c:5-119).  For ray data constant in s and phi, every
natural spline in s is the constant itself, so a plane's value is the tau sum of the
3-point Lagrange interpolant in z.  Its stencil is (iz, iz+1, iz+2) with
iz = trunc((z+1)/dz) (iz = ngrid-1 -> ngrid-2), and the third value is 0 in the top cell
(iz >= ngrid-2).  For data f(z_l):
  - f = c: the interpolant is c below the top cell and c (1 - l2(z)) in it, where
    l2(z) = (z - z0)(z - z1) / ((z2 - z0)(z2 - z1)) (a stencil clamped to the last three
    slices, or no zeroing, gives c there);
  - f = z^3: the interpolant is z^3 - (z - z0)(z - z1)(z - z2) (minus z2^3 l2(z) in the top
    cell); for z in [z0, z1] the remainder product is >= 0, while a centred stencil
    (iz-1, iz, iz+1) gives one <= 0, so the two differ in sign.
"""
import math

import numpy as np


def normals(phi, theta):
    """[nphi][ntheta][3] unit normals (sin t cos p, sin t sin p, cos t)"""
    p = np.asarray(phi)[:, None]
    t = np.asarray(theta)[None, :]
    return np.stack(np.broadcast_arrays(np.sin(t) * np.cos(p), np.sin(t) * np.sin(p),
                                        np.cos(t)), axis=-1)


def adjoint_axis_data(shift, phi, theta, k):
    """g[s][phi][theta] = s * n_k(phi, theta)"""
    nk = normals(phi, theta)[:, :, k]
    return np.asarray(shift)[:, None, None] * nk[None, :, :]


def voxel_centres(n):
    return -1.0 + 1.0 / n + np.arange(n) * (2.0 / n)


def adjoint_axis_expected(n, k):
    """(4 pi/3) x_k on the output grid out[z][x][y], and the mask of voxels with |x| <= 1"""
    c = voxel_centres(n)
    Z, X, Y = np.meshgrid(c, c, c, indexing="ij")
    coord = (X, Y, Z)[k]
    inside = X * X + Y * Y + Z * Z <= 1.0
    return 4.0 * math.pi / 3.0 * coord, inside


def _stencil(z, nz, centred=False):
    dz = 2.0 / (nz - 1)
    iz = int((z + 1.0) / dz)
    if iz == nz - 1:
        iz -= 1
    if centred:
        iz -= 1
    nodes = [(-1.0) + dz * (iz + m) for m in range(3)]
    return iz, nodes


def _interp_const(c, z, nz, zero_top=True):
    iz, (z0, z1, z2) = _stencil(z, nz)
    if zero_top and iz >= nz - 2:
        return c * (1.0 - (z - z0) * (z - z1) / ((z2 - z0) * (z2 - z1)))
    return c


def _interp_cubic(z, nz, centred=False):
    iz, (z0, z1, z2) = _stencil(z, nz, centred)
    v = z ** 3 - (z - z0) * (z - z1) * (z - z2)
    if iz >= nz - 2:
        v -= z2 ** 3 * (z - z0) * (z - z1) / ((z2 - z0) * (z2 - z1))
    return v


def _tau_sum(value, p, theta, nz):
    """sum over the reference's tau lattice of value(z) dtau where the plane point is inside
    (slicing.c:102-117): z = p cos t + tau sin t, s' = p sin t - tau cos t"""
    dtau = (2.0 / (nz - 1)) / 2.0
    ct, st = math.cos(theta), math.sin(theta)
    total = 0.0
    for i in range(-2 * nz, 2 * nz + 1):
        tau = i * dtau
        z = p * ct + tau * st
        s2 = p * st - tau * ct
        if abs(tau * tau + p * p) > 1.0 or abs(z) > 1.0 or abs(s2) > 1.0:
            continue
        total += value(z) * dtau
    return total


def reduction_constant(c, shift, theta, nz, zero_top=True):
    """[s][theta] plane values of ray data identically c (every phi alike)"""
    return np.array([[_tau_sum(lambda z: _interp_const(c, z, nz, zero_top), p, t, nz)
                      for t in theta] for p in shift])


def reduction_cubic(shift, theta, nz, centred=False):
    """[s][theta] plane values of ray data z_l^3 (every s and phi alike)"""
    return np.array([[_tau_sum(lambda z: _interp_cubic(z, nz, centred), p, t, nz)
                      for t in theta] for p in shift])


def constant_rays(nz, ns, nphi, c):
    return np.full((nz, ns, nphi), float(c))


def cubic_rays(nz, ns, nphi):
    z = (-1.0) + np.arange(nz) * (2.0 / (nz - 1))  # the slices z_l (ray_grid.c:244-247)
    return np.broadcast_to((z ** 3)[:, None, None], (nz, ns, nphi)).copy()
