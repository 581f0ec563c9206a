"""Signal helpers of the solver's post-processing.

``dtft`` restates rfnetwork/utils.py:14-71: X(f) = sum_n x[n] exp(-j 2 pi (f/fs) n) in float64,
optionally after a 20th-order Butterworth low-pass (second-order sections, forward-backward) and
decimation to 20x the highest frequency.
"""
import functools

import numpy as np
from scipy import signal


class LabeledArray(np.ndarray):
    """ndarray carrying a ``coords`` dict (the reference returns np_struct.ldarray objects)."""

    def __new__(cls, data, coords=None):
        obj = np.asarray(data).view(cls)
        obj.coords = dict(coords or {})
        return obj

    def __array_finalize__(self, obj):
        self.coords = getattr(obj, "coords", {})


_PHASE_CACHE_ENTRIES = 1 << 22  # tables up to 64 MiB (complex128) are cached, at most 8 of them


@functools.lru_cache(maxsize=8)
def _phases(n_samples: int, omega_bytes: bytes) -> np.ndarray:
    """exp(-j omega n) for n = 0 .. n_samples - 1 (rows) and each omega (columns), read-only.  The
    S-parameters of one solve transform every port and loop series over the same samples and
    frequencies, so the table is built once for all of them; the values are those computed in
    place, so the sums are unchanged."""
    omega = np.frombuffer(omega_bytes, dtype=np.float64)
    n = np.arange(n_samples)
    e = np.exp(-1j * omega[None, :] * n[:, None])
    e.flags.writeable = False
    return e


def dtft(xn: np.ndarray, frequency: np.ndarray, fs: float, downsample: bool = False) -> np.ndarray:
    """Discrete-time Fourier transform of x[n] sampled at fs, evaluated at ``frequency`` (Hz)."""
    xn = np.asarray(xn)
    frequency = np.atleast_1d(np.asarray(frequency, dtype=float))
    frs = np.max(frequency) * 20
    factor = int(fs / frs)
    use_ds = downsample and (len(xn) / factor) > 100 and fs > (frs * 2)
    if use_ds:
        sos = signal.butter(20, frs / 2.5, btype="lowpass", output="sos", fs=fs)
        xn = signal.sosfiltfilt(sos, xn)[::factor]
        fs = frs
    omega = 2 * np.pi * (frequency / fs)
    # one row of phases per sample; float64 throughout
    if len(xn) * len(omega) <= _PHASE_CACHE_ENTRIES:
        e = _phases(len(xn), omega.tobytes())
    else:
        e = np.exp(-1j * omega[None, :] * np.arange(len(xn))[:, None])
    Xw = np.sum(xn[:, None] * e, axis=0)
    if use_ds:
        Xw *= factor
    return Xw


def coupled_sline_impedance(w: float, s: float, b: float, er: float, t: float = 0):
    """Odd- and even-mode impedance (ohm) of edge-coupled stripline: strips of width w and thickness t,
    gap s, ground planes b apart, relative permittivity er (any one length unit).  The closed form of
    Pramanick and Bhartia, Modern RF and Microwave Filter Design, sec. 3.7.1, with the fringing factor
    Cf of a strip of thickness t; the same helper rfnetwork's coupled-line tests check against
    (rfnetwork/utils.py:199).  Returns (Z0_odd, Z0_even)."""
    ln2 = np.log(2)
    a_even = (ln2 + np.log(1 + np.tanh(np.pi * s / (2 * b)))) / (2 * np.pi * ln2)
    a_odd = (ln2 + np.log(1 + 1 / np.tanh(np.pi * s / (2 * b)))) / (2 * np.pi * ln2)
    cf = 2 * np.log((2 * b - t) / (b - t)) - (t / b) * (t * (2 * b - t) / (b - t) ** 2)
    z_even = 30 * np.pi * (b - t) / (np.sqrt(er) * (w + a_even * b * cf))
    z_odd = 30 * np.pi * (b - t) / (np.sqrt(er) * (w + a_odd * b * cf))
    return z_odd, z_even


def ustrip_impedance(w: float, h: float, er: float, frequency=1e9, t: float = 0.0013):
    """Characteristic impedance (ohm) and effective permittivity of a microstrip of width w and
    thickness t on a substrate h thick (inches) of relative permittivity er, at ``frequency`` (Hz) --
    what rfnetwork's microstrip element reports (rfnetwork/utils.py:295-365, elements.py:654-690) and
    its microstrip solver test checks against.

    Static part: Hammerstad and Jensen, "Accurate models for microstrip computer-aided design", IEEE
    MTT-S 1980 -- the air-line impedance Z01(u) = eta0/(2 pi) ln(F(u)/u + sqrt(1 + 4/u^2)), the
    effective permittivity eps_e(u) with its exponent a(u) b(er), and the finite-thickness width
    corrections (u1 for the air line, ur for the dielectric-filled one).  Dispersion: Kirschning and
    Jansen, Electronics Letters 18(6), 1982, with the normalised frequency f h in GHz mm; the impedance
    is scaled by sqrt(e / e_f) (e_f - 1) / (e - 1) as rfnetwork does.  Returns (Z0, eps_eff), scalars
    for a scalar frequency."""
    scalar = np.isscalar(frequency)
    f = np.atleast_1d(np.asarray(frequency, dtype=np.float64))
    eta0 = 376.73
    u, tn = w / h, t / h

    def z_air(x):
        F = 6 + (2 * np.pi - 6) * np.exp(-((30.666 / x) ** 0.7528))
        return eta0 / (2 * np.pi) * np.log(F / x + np.sqrt(1 + 4 / x ** 2))

    du1 = tn / np.pi * np.log(1 + 4 * np.e / tn * np.tanh(np.sqrt(6.517 * u)) ** 2)
    u1 = u + du1
    ur = u + du1 * (1 + 1 / np.cosh(np.sqrt(er - 1))) / 2
    a = 1 + np.log((ur ** 4 + (ur / 52) ** 2) / (ur ** 4 + 0.432)) / 49 + np.log(1 + (ur / 18.1) ** 3) / 18.7
    b = 0.564 * ((er - 0.9) / (er + 3)) ** 0.053
    eps_e = (er + 1) / 2 + (er - 1) / 2 * (1 + 10 / ur) ** (-a * b)
    z0 = z_air(ur) / np.sqrt(eps_e)
    e_static = eps_e * (z_air(u1) / z_air(ur)) ** 2

    fn = f / 1e9 * h * 25.4  # GHz * mm
    p1 = 0.27488 + (0.6315 + 0.525 / (1 + 0.0157 * fn) ** 20) * u - 0.065683 * np.exp(-8.7513 * u)
    p2 = 0.33622 * (1 - np.exp(-0.03442 * er))
    p3 = 0.0363 * np.exp(-4.6 * u) * (1 - np.exp(-((fn / 38.7) ** 4.97)))
    p4 = 1 + 2.751 * (1 - np.exp(-((er / 15.916) ** 8)))
    p = p1 * p2 * ((0.1844 + p3 * p4) * fn) ** 1.5763
    e_f = er - (er - e_static) / (1 + p)
    z_f = z0 * np.sqrt(e_static / e_f) * (e_f - 1) / (e_static - 1)
    if scalar:
        return z_f.item(), e_f.item()
    return z_f, e_f
