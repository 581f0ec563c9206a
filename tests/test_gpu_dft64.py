"""Running DFTs accumulated in double in a float32 solve (rfdtd_monitor.flags = RFDTD_MONITOR_DFT64):
each sampled float32 plane value is widened exactly and multiplied by the complex128 phase entry,
and the products are added in time order.  The expected sums are computed here in numpy from the
run's own time-domain samples of the same plane, with the same float64 operations in the same order,
so the comparison is exact; the complex64 monitor of the same plane is checked the same way in
float32 arithmetic."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, hip_func, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

pytestmark = pytest.mark.gpu

FREQS = np.array([10e9, 15e9, 20e9, 30e9, 35e9])  # 5 frequencies: one full group of 4 and a partial one


def _model(Nt, seed=7):
    m = workloads.tiny_model(Nt=Nt, seed=seed, time_monitors=False, phasor_monitors=False)
    m.add_field_monitor("t_ez", "ez", "x", index=8, n_step=1)
    m.add_field_monitor("p_ez32", "ez", "x", index=8, frequency=FREQS)
    m.add_field_monitor("p_ez64", "ez", "x", index=8, frequency=FREQS, dft_dtype=np.complex128)
    m.add_field_monitor("t_hx", "hx", "y", index=6, n_step=2)
    m.add_field_monitor("p_hx64", "hx", "y", index=6, frequency=FREQS, n_step=2, dft_dtype=np.complex128)
    return m


def _dft_in_time_order(samples, phase, n_step, Nt, real):
    """sum_m samples[m] * phase[m, f] over the sampled rows m = n / n_step, n = 0 .. Nt-1, one product
    and one sum per component in `real` arithmetic, in time order."""
    rows = (Nt - 1) // n_step + 1
    nf = phase.shape[1]
    re = np.zeros((nf,) + samples.shape[1:], real)
    im = np.zeros_like(re)
    for m in range(rows):
        v = samples[m].astype(real)[None]
        re = re + v * phase[m].real.astype(real)[:, None, None]
        im = im + v * phase[m].imag.astype(real)[:, None, None]
    return re, im


def _check(monitors, Nt):
    t_ez, p32, p64, t_hx, q64 = (np.asarray(m["values"]) for m in monitors)
    assert p64.dtype == np.complex128 and q64.dtype == np.complex128 and p32.dtype == np.complex64
    re, im = _dft_in_time_order(t_ez, monitors[2]["dtft_phase"], 1, Nt, np.float64)
    np.testing.assert_array_equal(p64.real, re)
    np.testing.assert_array_equal(p64.imag, im)
    re, im = _dft_in_time_order(t_ez, monitors[1]["dtft_phase"], 1, Nt, np.float32)
    np.testing.assert_array_equal(p32.real, re)
    np.testing.assert_array_equal(p32.imag, im)
    re, im = _dft_in_time_order(t_hx, monitors[4]["dtft_phase"], 2, Nt, np.float64)
    np.testing.assert_array_equal(q64.real, re)
    np.testing.assert_array_equal(q64.imag, im)
    # the two precisions agree to float32 rounding of the sums
    scale = np.abs(p64).max()
    assert scale > 0
    np.testing.assert_allclose(p32, p64, rtol=0, atol=1e-5 * scale)


@pytest.mark.parametrize("variant,flags", [(1, 0), (4, _lib.NO_GRAPHS), (6, 0), (0, _lib.COEF_RAW)])
def test_dft64_equals_float64_sum_of_the_samples(hip_ready, variant, flags):
    print(f"[case] dft64 variant={variant} flags={flags}", flush=True)
    Nt = 57
    model = _model(Nt)
    c, probes, monitors, _, Nx, Ny, Nz, _ = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz, flags=flags) as s:
        s.set_variant(variant)
        s.set_sources(probes, monitors, Nt)
        for n in (23, 1, 33):  # flushes of full and partial staging sets
            s.run(n)
        s.synchronize()
        s.read()
    _check(monitors, Nt)


def test_dft64_drop_in_and_batch(hip_ready):
    """The drop-in entry point and a batch of two excitations give the same double sums as a
    session."""
    print("[case] dft64 drop-in and batch", flush=True)
    Nt = 41
    model = _model(Nt)
    c, probes, monitors, mem, Nx, Ny, Nz, _ = model.pack(gpu=True)
    hip_func.solver_run_hip(c, probes, monitors, mem, Nx, Ny, Nz, Nt, 4, 0)
    _check(monitors, Nt)
    models = [_model(Nt, seed=7), _model(Nt, seed=9)]
    grids, inputs = [], []
    for m in models:
        grids.append((m.coefficient_arrays(gpu=True), m.Nx, m.Ny, m.Nz, 1))
        _, pr, mo, _, *_ = m.pack(gpu=True, coefficients=False)
        inputs.append((pr, mo))
    with Batch(grids) as b:
        for k, (pr, mo) in enumerate(inputs):
            b.set_sources(k, pr, mo, Nt)
        b.run(Nt)
        b.synchronize()
        b.read()
    for _, mo in inputs:
        _check(mo, Nt)
    for a, b_ in zip(monitors, inputs[0][1]):
        np.testing.assert_array_equal(np.asarray(a["values"]), np.asarray(b_["values"]))
