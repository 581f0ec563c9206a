"""rfdtd_nf2ff_static: the far-field integrals staged through the library's static device workspace
(no device allocation) give exactly what rfdtd_nf2ff gives on the same inputs, on two calls in a
row (the second smaller, so stale workspace contents from the first would show), and requests
beyond the 64 MiB workspace are refused before anything is copied."""
import numpy as np
import pytest

from rfnetwork_amd import _lib
from rfnetwork_amd.units import const

pytestmark = pytest.mark.gpu


def _faces(rng, spec, nf):
    faces = []
    for axis, pos, n0, n1 in spec:
        r0, r1 = np.meshgrid(np.linspace(-0.01, 0.012, n0), np.linspace(-0.008, 0.009, n1), indexing="ij")
        c = lambda: (rng.standard_normal((3, nf, n0, n1)) + 1j * rng.standard_normal((3, nf, n0, n1))).astype(np.complex64)
        faces.append(dict(axis=axis, position=pos, r0=r0.astype(np.float32), r1=r1.astype(np.float32),
                          ds=rng.uniform(0.5e-6, 2e-6, (n0, n1)).astype(np.float32), J=c(), M=c()))
    return faces


def test_static_workspace_equals_allocating_entry(hip_ready):
    print("[case] nf2ff_static two calls", flush=True)
    rng = np.random.default_rng(33)
    eta0 = float(const.eta0)
    big = _faces(rng, ((0, -0.011, 40, 31), (1, 0.013, 7, 300), (2, 0.004, 63, 45), (0, 0.012, 40, 31)), 3)
    beta1 = np.array([120.0, 260.0, 410.0], np.float32)
    th1 = np.deg2rad(np.linspace(0, 180, 37)).astype(np.float32)
    ph1 = np.deg2rad(np.linspace(0, 355, 72)).astype(np.float32)
    small = _faces(rng, ((2, -0.003, 9, 5),), 1)
    beta2 = np.array([333.0], np.float32)
    th2 = np.deg2rad(np.linspace(10, 170, 5)).astype(np.float32)
    ph2 = np.deg2rad(np.linspace(0, 90, 3)).astype(np.float32)
    want1 = _lib.nf2ff(big, beta1, th1, ph1, eta0)
    want2 = _lib.nf2ff(small, beta2, th2, ph2, eta0)
    got1 = _lib.nf2ff(big, beta1, th1, ph1, eta0, static_workspace=True)
    got2 = _lib.nf2ff(small, beta2, th2, ph2, eta0, static_workspace=True)
    assert np.abs(want1).max() > 0 and np.abs(want2).max() > 0
    np.testing.assert_array_equal(got1, want1)
    np.testing.assert_array_equal(got2, want2)
    # and once more in the other order: the large request after the small one
    np.testing.assert_array_equal(_lib.nf2ff(big, beta1, th1, ph1, eta0, static_workspace=True), want1)


def test_static_workspace_refuses_oversized_requests(hip_ready):
    print("[case] nf2ff_static size limit", flush=True)
    n0 = n1 = 1000  # J and M of 3 frequencies: 72 MB each, beyond the 64 MiB workspace
    face = dict(axis=2, position=0.0, r0=np.zeros((n0, n1), np.float32), r1=np.zeros((n0, n1), np.float32),
                ds=np.zeros((n0, n1), np.float32), J=np.zeros((3, 3, n0, n1), np.complex64),
                M=np.zeros((3, 3, n0, n1), np.complex64))
    with pytest.raises(ValueError, match="64 MiB"):
        _lib.nf2ff([face], np.ones(3, np.float32), np.zeros(2, np.float32), np.zeros(2, np.float32), 377.0,
                   static_workspace=True)
    with pytest.raises(ValueError, match="at most 64 faces"):
        small = dict(axis=2, position=0.0, r0=np.zeros((1, 1), np.float32), r1=np.zeros((1, 1), np.float32),
                     ds=np.zeros((1, 1), np.float32), J=np.zeros((3, 1, 1, 1), np.complex64),
                     M=np.zeros((3, 1, 1, 1), np.complex64))
        _lib.nf2ff([small] * 65, np.ones(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32), 377.0,
                   static_workspace=True)
