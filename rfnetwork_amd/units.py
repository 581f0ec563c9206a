"""Physical constants and unit conversions used by the FDTD solver.

The values match rfnetwork/core/units.py:19-29 (const) and :105-124 (conv): parity of the
coefficients depends on them, in particular on inches -> metres being a multiplication by
1/39.3701 (not 0.0254) and on c0 being the integer 299792458.
"""
import numpy as np


class const:
    """Physical constants (SI)."""

    c0 = 299792458
    e0 = 8.854187817e-12
    u0 = 4 * np.pi * 1e-7
    eta0 = np.sqrt(u0 / e0)
    k = 1.380649e-23
    t0 = 290


class conv:
    """Unit conversions named ``<to>_<from>``: ``conv.m_in(x)`` converts inches to metres."""

    ratios = dict(in_m=39.3701, mm_in=25.4, mil_mm=39.3701, ft_m=3.28084)

    @staticmethod
    def db10(x):
        return 10 * np.log10(np.abs(x))

    @staticmethod
    def db20(x):
        return 20 * np.log10(np.abs(x))

    @staticmethod
    def z_gamma(gamma, z0=50):
        """Impedance seen through a reflection coefficient."""
        return z0 * (1 + gamma) / (1 - gamma)

    @staticmethod
    def gamma_z(z, z0=50):
        return (z - z0) / (z + z0)


def _ratio_fn(r):
    return lambda v: r * v


for _k, _v in conv.ratios.items():
    _a, _b = _k.split("_")
    # same construction as the reference: <a>_<b>(v) = ratio * v and <b>_<a>(v) = (1/ratio) * v
    setattr(conv, f"{_a}_{_b}", staticmethod(_ratio_fn(_v)))
    setattr(conv, f"{_b}_{_a}", staticmethod(_ratio_fn(1 / _v)))
