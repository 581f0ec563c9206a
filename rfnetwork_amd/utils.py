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
