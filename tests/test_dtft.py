"""utils.dtft (restating rfnetwork/utils.py:14-71) against textbook identities."""
import numpy as np
from scipy import signal

from rfnetwork_amd import utils


def test_dtft_equals_fft_at_bins():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(256)
    fs = 1e12
    k = np.arange(0, 40)
    f = k * fs / len(x)
    np.testing.assert_allclose(utils.dtft(x, f, fs), np.fft.fft(x)[:40], rtol=1e-10, atol=1e-10)


def test_dtft_of_delayed_impulse():
    x = np.zeros(100)
    x[7] = 2.0
    f = np.linspace(1e9, 5e9, 9)
    fs = 50e9
    np.testing.assert_allclose(utils.dtft(x, f, fs), 2 * np.exp(-2j * np.pi * f / fs * 7), rtol=1e-12)


def test_downsampled_dtft_matches_full_for_band_limited_signal():
    # the decimated series is treated as sampled at exactly 20 x max(f) (rfnetwork/utils.py:44-50), which
    # holds when fs is an integer multiple of it
    fs = 62 * 20 * 8e9
    t = np.arange(20000) / fs
    x = np.exp(-((t - 0.5e-9) / 0.1e-9) ** 2) * np.sin(2 * np.pi * 5e9 * t)
    f = np.linspace(2e9, 8e9, 7)
    full = utils.dtft(x, f, fs, downsample=False)
    ds = utils.dtft(x, f, fs, downsample=True)
    assert np.max(np.abs(ds - full)) < 1e-3 * np.max(np.abs(full))
    # the decimation path is a 20th-order Butterworth low-pass applied forward and backward
    sos = signal.butter(20, np.max(f) * 20 / 2.5, btype="lowpass", output="sos", fs=fs)
    assert sos.shape == (10, 6)


def test_cached_phase_table_gives_the_same_bits(monkeypatch):
    """The phase table shared by the transforms of one solve changes no bit: a cached call, a
    repeated call and a call with caching switched off agree exactly, and the cached table cannot be
    written through."""
    rng = np.random.default_rng(5)
    f = np.linspace(1e9, 12e9, 57)
    x = rng.standard_normal(1500).astype(np.float32)
    utils._phases.cache_clear()
    first = utils.dtft(x, f, 3e11)
    again = utils.dtft(x, f, 3e11)
    assert utils._phases.cache_info().hits >= 1
    monkeypatch.setattr(utils, "_PHASE_CACHE_ENTRIES", 0)
    direct = utils.dtft(x, f, 3e11)
    assert np.array_equal(first, again) and np.array_equal(first, direct)
    table = utils._phases(len(x), (2 * np.pi * (f / 3e11)).tobytes())
    assert not table.flags.writeable
