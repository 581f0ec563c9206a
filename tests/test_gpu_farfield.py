"""Far field (NF2FF): the GPU radiation integrals against a float64 direct evaluation of the same sums
(rfnetwork/core/src/postprocess.cpp:130-505, Balanis 6.8.2), and rfnetwork's own half-wave dipole
test case (gain, polarisation, nulls, match) solved by the engine."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.units import const

pytestmark = pytest.mark.gpu


def _direct(faces, beta, theta, phi, eta0):
    """float64 restatement of the radiation integrals."""
    nf = len(beta)
    out = np.zeros((2, nf, len(theta), len(phi)), complex)
    for it, th in enumerate(theta):
        for ip, ph in enumerate(phi):
            ct, st, cp, sp = np.cos(th), np.sin(th), np.cos(ph), np.sin(ph)
            u, v, w = st * cp, st * sp, ct
            for fi, b in enumerate(beta):
                Nt = Np = Lt = Lp = 0j
                for F in faces:
                    r0, r1 = F["r0"].astype(float), F["r1"].astype(float)
                    pos = F["position"]
                    rdot = {0: pos * u + r0 * v + r1 * w, 1: pos * v + r0 * u + r1 * w, 2: pos * w + r0 * u + r1 * v}[F["axis"]]
                    I = np.exp(1j * b * rdot) * F["ds"].astype(float)
                    J, M = F["J"][:, fi].astype(complex), F["M"][:, fi].astype(complex)
                    Nt += np.sum((J[0] * ct * cp + J[1] * ct * sp - J[2] * st) * I)
                    Np += np.sum((-J[0] * sp + J[1] * cp) * I)
                    Lt += np.sum((M[0] * ct * cp + M[1] * ct * sp - M[2] * st) * I)
                    Lp += np.sum((-M[0] * sp + M[1] * cp) * I)
                out[0, fi, it, ip] = (-1j * b / (4 * np.pi)) * (Lp + eta0 * Nt)
                out[1, fi, it, ip] = (1j * b / (4 * np.pi)) * (Lt - eta0 * Np)
    return out


def test_nf2ff_kernel_matches_direct_sum(hip_ready):
    print("[case] nf2ff kernel", flush=True)
    rng = np.random.default_rng(21)
    faces = []
    for axis, pos, n0, n1 in ((0, -0.011, 17, 9), (1, 0.013, 5, 300), (2, 0.004, 23, 31)):
        r0, r1 = np.meshgrid(np.linspace(-0.01, 0.012, n0), np.linspace(-0.008, 0.009, n1), indexing="ij")
        c = lambda: (rng.standard_normal((3, 2, n0, n1)) + 1j * rng.standard_normal((3, 2, n0, n1))).astype(np.complex64)
        faces.append(dict(axis=axis, position=pos, r0=r0.astype(np.float32), r1=r1.astype(np.float32),
                          ds=rng.uniform(0.5e-6, 2e-6, (n0, n1)).astype(np.float32), J=c(), M=c()))
    beta = np.array([120.0, 410.0], np.float32)
    theta = np.deg2rad(np.linspace(0, 180, 13)).astype(np.float32)
    phi = np.deg2rad(np.linspace(0, 355, 7)).astype(np.float32)
    got = _lib.nf2ff(faces, beta, theta, phi, float(const.eta0))
    want = _direct(faces, beta.astype(float), theta.astype(float), phi.astype(float), float(const.eta0))
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[nf2ff] max error / max value {err:.2e}", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("mesh", [(0.0075, 0.0075, 1.0), (0.01, 0.03, 1.2)])
def test_half_wave_dipole_known_answers(hip_ready, mesh):
    """rfnetwork's own dipole test (tests/test_solver/test_dipole.py:100-125) on the engine, with the
    reference's dimensions on a grid through the strip edges, gap and arm ends (0.0075 inch cells, and
    the reference mesher's d_min 0.01 / d_max 0.03 graded by 1.2 per cell): realised theta-polarised
    gain at theta = 90 deg within 0.35 dB of 2.1 dB and below 2.15 dB for every phi, phi-polarised gain
    below -40 dB, nulls below -90 dB at the poles, S11 below -10 dB and the input resistance within
    3.5 ohm of 73 ohm: the reference's bounds.  Measured: 74.65 + 5.42j ohm, S11 -13.9 dB, gain
    1.876-2.082 dB, phi-pol -52 dB on the first grid; 75.96 + 8.82j ohm, gain 1.763-2.133 dB, phi-pol
    -40.5 dB on the graded one (finer gradations put the gain up to 2.24 dB at some phi: the E and H
    averages to the face cell centres are unweighted, as in the reference)."""
    print(f"[case] half-wave dipole {mesh}", flush=True)
    s = workloads.half_wave_dipole_exact(*mesh)
    s.solve(gpu=True, show_progress=False)
    gain = np.asarray(s.get_farfield_gain(theta=np.arange(0, 190, 10), phi=np.arange(-180, 184, 4)))
    gdb = 10 * np.log10(gain)
    g90 = gdb[0, 0, 9]  # thetapol, 10 GHz, theta = 90
    print(f"[farfield] thetapol gain at 90 deg: {g90.min():.3f} .. {g90.max():.3f} dB", flush=True)
    assert np.all(g90 < 2.15)
    assert np.all(np.abs(2.1 - g90) < 0.35)
    assert np.all(gdb[1] < -40)
    assert np.all(gdb[0, :, 0] < -90) and np.all(gdb[0, :, -1] < -90)
    s11 = np.asarray(s.get_sparameters([10e9]))[:, 0, 0]
    assert np.all(20 * np.log10(np.abs(s11)) < -10)
    z = 50 * (1 + s11) / (1 - s11)
    print(f"[farfield] S11 {20 * np.log10(np.abs(s11))} dB, Z {z} ohm", flush=True)
    assert np.all(np.abs(73 - z.real) < 3.5)
