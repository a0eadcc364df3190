"""Analytic test fields of the reference (test_fields/test_fields_mod.f90), restated in numpy.

TEST INFRASTRUCTURE ONLY.  Every generator takes Cartesian point coordinates
(x, y, z) of mesh points and returns a scalar or a Cartesian vector, as the
reference's get_scalar_field / get_vector_field do; `level` selects the level
index k (1-based) for the generators whose definition cycles over k.
"""
import numpy as np


def xyz_scalar(x, y, z, level=1):
    """generate_xyz_scalar_field, test_fields_mod.f90:463-480: level 1 -> x, 2 -> y, 3 -> z."""
    k1 = (level - 1) % 3
    M = np.eye(3)
    return M[0, k1] * x + M[1, k1] * y + M[2, k1] * z


def solid_rotation(x, y, z, level=1):
    """generate_solid_rotation_vector_field, test_fields_mod.f90:505-524 (axis e_k per level)."""
    k1 = (level - 1) % 3
    ax = np.eye(3)[:, k1]
    vx = ax[1] * z - ax[2] * y
    vy = -ax[0] * z + ax[2] * x
    vz = ax[0] * y - ax[1] * x
    return vx, vy, vz


def xyz_grad(x, y, z, level=1):
    """generate_xyz_grad_field, test_fields_mod.f90:631-660: horizontal gradient of x, y or z."""
    k1 = (level - 1) % 3
    e = np.eye(3)[:, k1]
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    n = (x / r, y / r, z / r)
    ne = n[0] * e[0] + n[1] * e[1] + n[2] * e[2]
    return e[0] - n[0] * ne, e[1] - n[1] * ne, e[2] - n[2] * ne


def _sin_phi(x, y, z):
    return z / np.sqrt(x ** 2 + y ** 2 + z ** 2)


def _cos_phi(x, y, z):
    return np.sqrt((x ** 2 + y ** 2) / (x ** 2 + y ** 2 + z ** 2))


def _cos_lam(x, y, z):
    return x / np.maximum(np.sqrt(x ** 2 + y ** 2), 1e-14)


def _sin_lam(x, y, z):
    s = y / np.maximum(np.sqrt(x ** 2 + y ** 2), 1e-14)
    return np.where(x ** 2 + y ** 2 == 0, 1.0, s)


def cross_polar(x, y, z, level=1):
    """generate_cross_polar_flow, test_fields_mod.f90:662-701 (latlon_functions_mod.f90:4-26)."""
    cphi, sphi = _cos_phi(x, y, z), _sin_phi(x, y, z)
    clam, slam = _cos_lam(x, y, z), _sin_lam(x, y, z)
    u = sphi * (sphi ** 2 - 3.0 * cphi ** 2) * slam - 0.5 * cphi
    v = sphi ** 2 * clam
    return -slam * u - sphi * clam * v, clam * u - sphi * slam * v, cphi * v


def cross_polar_div(x, y, z, level=1):
    """generate_cross_polar_flow_div, test_fields_mod.f90:703-716."""
    return -x * z / (x ** 2 + y ** 2 + z ** 2)


def cross_polar_zeta(x, y, z, level=1):
    """generate_cross_polar_flow_zeta, test_fields_mod.f90:718-731."""
    return (16.0 * (1 - z ** 2) * y - 13.0 * y - z) / (x ** 2 + y ** 2 + z ** 2)


def _cart2sph(x, y, z):
    """sph_coords_mod.f90:20-30."""
    lam = np.arctan2(y, x)
    phi = np.arctan2(z, np.sqrt(x ** 2 + y ** 2))
    lam = lam + np.pi * (1.0 - np.where(lam >= 0, 1.0, -1.0))
    return lam, phi


def _sph2cart_vec(lam, phi, vl, vp):
    """sph_coords_mod.f90:32-41."""
    vx = -np.sin(lam) * vl - np.sin(phi) * np.cos(lam) * vp
    vy = np.cos(lam) * vl - np.sin(phi) * np.sin(lam) * vp
    vz = np.cos(phi) * vp
    return vx, vy, vz


def vsh_curl_free_10(x, y, z, level=1):
    """generate_VSH_curl_free_10, test_fields_mod.f90:755-774: v_phi = -cos(phi)."""
    lam, phi = _cart2sph(x, y, z)
    return _sph2cart_vec(lam, phi, 0.0 * lam, -np.cos(phi))


def _cart2sph_vec(lam, phi, vx, vy, vz):
    """sph_coords_mod.f90:43-51."""
    vl = -np.sin(lam) * vx + np.cos(lam) * vy
    vp = -np.cos(lam) * np.sin(phi) * vx - np.sin(lam) * np.sin(phi) * vy + np.cos(phi) * vz
    return vl, vp


def coriolis_force(field):
    """coriolis_force_field_generator_t, test_fields_mod.f90:592-616: f x V with f = 2 sin(phi)
    (unit rotation rate) of the vector field `field`."""
    def gen(x, y, z, level=1):
        vx, vy, vz = field(x, y, z, level)
        lam, phi = _cart2sph(x, y, z)
        vl, vp = _cart2sph_vec(lam, phi, vx, vy, vz)
        pc = 2 * np.sin(phi)
        return _sph2cart_vec(lam, phi, pc * vp, -pc * vl)
    return gen


def kinetic_energy(field):
    """KE_scalar_field_t, test_fields_mod.f90:804-825: 0.5 |V|^2 of the vector field `field`."""
    def gen(x, y, z, level=1):
        vx, vy, vz = field(x, y, z, level)
        return 0.5 * (vx ** 2 + vy ** 2 + vz ** 2)
    return gen
