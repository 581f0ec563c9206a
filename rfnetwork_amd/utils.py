"""Signal helpers of the solver's post-processing.

``dtft`` restates rfnetwork/utils.py:14-71: X(f) = sum_n x[n] exp(-j 2 pi (f/fs) n) in float64,
optionally after a 20th-order Butterworth low-pass (second-order sections, forward-backward) and
decimation to 20x the highest frequency.
"""
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
    n = np.arange(len(xn))
    # one row of phases per sample; float64 throughout
    Xw = np.sum(xn[:, None] * np.exp(-1j * omega[None, :] * n[:, None]), axis=0)
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
