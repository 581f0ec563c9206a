"""HIP engine vs the CPU oracle (restated rfnetwork/core/src/solver.cpp) on the same inputs.

Float32 and float64 results must be identical bit for bit: both evaluate the same per-element
expressions with separately rounded products and sums (no fused multiply-add).
"""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.hip_func import solver_run_hip
from rfnetwork_amd.session import Session

from helpers import graded_everything as _graded_everything, rel_l2, run_oracle

pytestmark = pytest.mark.gpu


def _hip_session(model, dtype=np.float32, flags=0, chunks=None, launch=None, variant=None):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype)
    with Session(c, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        if launch:
            s.set_launch(*launch)
        if variant is not None:
            s.set_variant(variant)
        s.set_sources(probes, monitors, Nt)
        for k in chunks or [Nt]:
            s.run(k)
        s.synchronize()
        s.read()
        f = s.fields()
        info = s.memory_info()
    return probes, monitors, f, info


def _global_monitors(model, probes, monitors, gpu):
    """Monitor planes padded to full Yee planes by FDTD_Solver.unpack (both packings)."""
    model.unpack(probes, monitors, gpu=gpu)
    return {k: np.array(m["values"]) for k, m in model.monitors.items()}


def _assert_monitors_equal(model, p_g, m_g, p_c, m_c):
    g = _global_monitors(model, p_g, m_g, True)
    c = _global_monitors(model, p_c, m_c, False)
    for k in g:
        assert g[k].shape == c[k].shape, k
        np.testing.assert_array_equal(g[k], c[k], err_msg=k)


@pytest.mark.parametrize("flags", [0, _lib.COEF_RAW, _lib.NO_GRAPHS, _lib.COEF_TABLE | _lib.NO_GRAPHS])
def test_tiny_fp32_bit_exact(hip_ready, flags):
    model = workloads.tiny_model(Nt=60)
    print(f"[case] tiny fp32 flags={flags}", flush=True)
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=3)
    p_g, m_g, f_g, info = _hip_session(model, np.float32, flags, chunks=[37, 23], launch=(64, 3, 10))
    for a, b in zip(p_g, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    _assert_monitors_equal(model, p_g, m_g, p_c, m_c)
    for k in f_c:
        np.testing.assert_array_equal(f_g[k], f_c[k], err_msg=k)
    want = "raw" if flags & _lib.COEF_RAW else "table"
    assert info["coef_kind"] == want


def test_tiny_fp64_bit_exact(hip_ready):
    model = workloads.tiny_model(Nt=50)
    print("[case] tiny fp64", flush=True)
    p_c, m_c, f_c = run_oracle(model, np.float64, n_threads=2)
    p_g, m_g, f_g, _ = _hip_session(model, np.float64)
    for a, b in zip(p_g, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    _assert_monitors_equal(model, p_g, m_g, p_c, m_c)
    for k in f_c:
        np.testing.assert_array_equal(f_g[k], f_c[k], err_msg=k)


def test_launch_shapes_agree(hip_ready):
    """Block size, x planes per block and graph chunk length do not change a single bit."""
    model = workloads.tiny_model(Nt=40, nx=21, ny=13, nz=70)
    print("[case] launch shapes", flush=True)
    ref = None
    for launch in [(256, 0, 100), (64, 1, 7), (1024, 5, 40), (128, 21, 1)]:
        for variant in (1, 0, 4):
            _, _, f, _ = _hip_session(model, launch=launch, variant=variant)
            if ref is None:
                ref = f
            else:
                for k in ref:
                    np.testing.assert_array_equal(f[k], ref[k], err_msg=f"{k} {launch} {variant}")


@pytest.mark.parametrize("dtype,flags", [(np.float32, _lib.COEF_RAW), (np.float32, _lib.COEF_TABLE),
                                         (np.float64, _lib.COEF_RAW)])
def test_vector_columns_match_oracle(hip_ready, dtype, flags):
    """The separate kernels over 2 and 4 consecutive z cells per thread (variants 11 and 12, Nz a
    multiple of 4): probes, monitors and every field bit-exact against the oracle, across launch
    shapes and graph chunk lengths (a group reaching the row's end, one starting a row, x chunks of
    1 .. all planes)."""
    model = workloads.tiny_model(Nt=60, nx=15, ny=11, nz=20)
    print(f"[case] vector columns dtype={np.dtype(dtype).name} flags={flags}", flush=True)
    p_c, m_c, f_c = run_oracle(model, dtype, n_threads=3)
    for variant in (11, 12):
        for launch in [(256, 0, 100), (64, 1, 7), (128, 15, 1), (512, 4, 60)]:
            p_g, m_g, f_g, _ = _hip_session(model, dtype, flags, chunks=[29, 31], launch=launch, variant=variant)
            for a, b in zip(p_g, p_c):
                np.testing.assert_array_equal(a["values"], b["values"], err_msg=f"{variant} {launch}")
            _assert_monitors_equal(model, p_g, m_g, p_c, m_c)
            for k in f_c:
                np.testing.assert_array_equal(f_g[k], f_c[k], err_msg=f"{k} {variant} {launch}")


@pytest.mark.parametrize("dtype,flags", [(np.float32, 0), (np.float32, _lib.COEF_RAW), (np.float64, 0)])
def test_column_order_launch_order_matches_oracle(hip_ready, dtype, flags):
    """The separate kernels with workgroups in launch order (rfdtd_set_column_order 1) instead of XCD
    ranges: every separate variant bit-exact against the oracle on the graded grid, across launch
    shapes, and the order reads back."""
    print(f"[case] column order dtype={np.dtype(dtype).name} flags={flags}", flush=True)
    model = _graded_everything(Nt=50, coef_dtype=dtype)
    p_c, m_c, f_c = run_oracle(model, dtype, n_threads=3)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype)
    for variant in (0, 1, 4, 10, 11, 12):
        for launch in [(256, 1, 100), (128, 3, 7), (512, 0, 1)]:
            _, probes, monitors, _, *_ = model.pack(gpu=True, dtype=dtype, coefficients=False)
            with Session(c, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
                s.set_launch(*launch)
                s.set_variant(variant)
                s.set_column_order(1)
                assert s.column_order() == 1
                s.set_sources(probes, monitors, Nt)
                s.run(29)
                s.run(Nt - 29)
                s.synchronize()
                s.read()
                f = s.fields()
            for a, b in zip(probes, p_c):
                np.testing.assert_array_equal(a["values"], b["values"], err_msg=f"{variant} {launch}")
            _assert_monitors_equal(model, probes, monitors, p_c, m_c)
            for k in f_c:
                np.testing.assert_array_equal(f[k], f_c[k], err_msg=f"{k} {variant} {launch}")
    with Session(c, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        with pytest.raises(ValueError):
            s.set_column_order(2)


def test_vector_columns_need_nz_multiple(hip_ready):
    """Variants 11 and 12 are refused on a grid whose Nz is not a multiple of 2 or 4."""
    model = workloads.tiny_model(Nt=10, nz=9)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        for v in (11, 12):
            with pytest.raises(ValueError, match="multiple"):
                s.set_variant(v)


def test_solve_api_matches_oracle(hip_ready):
    """FDTD_Solver.solve(gpu=True) end to end: port series, probes, monitors and S-parameters."""
    from oracle import cpu_solver

    m_g = workloads.tiny_model(Nt=80)
    m_c = workloads.tiny_model(Nt=80)
    print("[case] solve api", flush=True)
    m_g.solve(gpu=True, show_progress=False)
    m_c.solve(gpu=False, backend=cpu_solver.solver_run, n_threads=2, show_progress=False)
    np.testing.assert_array_equal(m_g.vi_probe_values("v1"), m_c.vi_probe_values("v1"))
    for p_g, p_c in zip(m_g.ports, m_c.ports):
        np.testing.assert_array_equal(p_g["values"], p_c["values"])
    S_g = m_g.get_sparameters(m_g.frequency)
    S_c = m_c.get_sparameters(m_c.frequency)
    np.testing.assert_array_equal(S_g, S_c)
    for k in m_g.monitors:
        a, b = m_g.monitors[k]["values"], m_c.monitors[k]["values"]
        assert a.shape == b.shape, k
        np.testing.assert_array_equal(a, b, err_msg=k)


def test_shared_cell_probes_in_list_order(hip_ready):
    """Several probes on one cell (a source and plain probes before and after it) are applied in list
    order, as the CPU solver's per-slab loop does."""
    model = workloads.tiny_model(Nt=30, time_monitors=False, phasor_monitors=False)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    src = probes[0]
    extra = [dict(values=np.zeros(Nt, np.float32), field=src["field"], idx=list(src["idx"]), is_source=0),
             dict(values=np.linspace(0, 1e-3, Nt).astype(np.float32), field=src["field"], idx=list(src["idx"]),
                  is_source=1)]
    probes = [extra[0]] + probes + [extra[1]]
    cc, pc, mc, memc, *_ = model.pack(gpu=False)
    pc = [dict(values=extra[0]["values"].copy(), field=extra[0]["field"], idx=extra[0]["idx"], is_source=0)] + pc
    pc += [dict(values=extra[1]["values"].copy(), field=extra[1]["field"], idx=extra[1]["idx"], is_source=1)]
    print("[case] shared-cell probes", flush=True)
    from oracle import cpu_solver
    cpu_solver.solver_run(cc, pc, mc, memc, Nx, Ny, Nz, Nt, 2, 0)
    solver_run_hip(c, probes, monitors, mem, Nx, Ny, Nz, Nt, 0, 0)
    for a, b in zip(probes, pc):
        np.testing.assert_array_equal(a["values"], b["values"])


def test_microstrip_4000_steps(hip_ready):
    """The 120x48x32 microstrip, 4000 steps: fields, port series and S-parameters against the oracle,
    and the float32 error measured against a float64 run of the same kernels."""
    print("[case] microstrip 4000 steps", flush=True)
    model = workloads.microstrip()
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=8)
    p_g, m_g, f_g, info = _hip_session(model, np.float32)
    errs = {k: rel_l2(f_g[k], f_c[k]) for k in f_c}
    assert max(errs.values()) <= 1e-5, errs
    for a, b in zip(p_g, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    # S-parameters through the model API
    model.unpack(p_g, m_g, gpu=True)
    S_g = np.array(model.get_sparameters(model.frequency))
    model.unpack(p_c, m_c, gpu=False)
    S_c = np.array(model.get_sparameters(model.frequency))
    assert np.max(np.abs(np.abs(S_g) - np.abs(S_c))) <= 1e-4
    assert np.max(np.abs(np.angle(S_g / S_c, deg=True))) <= 0.05
    # float32 vs float64 run of the same kernels (stated, bounded loosely)
    p_d, m_d, f_d, _ = _hip_session(model, np.float64)
    e64 = max(rel_l2(f_g[k], f_d[k]) for k in ("ex", "ey", "ez", "hx", "hy", "hz"))
    model.unpack(p_d, m_d, gpu=True)
    S_d = np.array(model.get_sparameters(model.frequency))
    s64 = np.max(np.abs(S_g - S_d))
    print(f"[microstrip] fp32-vs-fp64 field rel L2 {e64:.3e}, max |dS| {s64:.3e}, coef {info}", flush=True)
    # after 4000 steps the pulse has left through the ports and PML; what remains is small residual
    # field, so its relative L2 drift is large (measured 5.1e-3) while S agrees to ~5e-7
    assert e64 < 5e-2 and s64 < 1e-5


def test_large_grid_parity_and_linearity(hip_ready):
    """Full 400x400x200 grid: 440 steps against the oracle -- fields, probes and the 6 running-DFT
    planes bit for bit, at a step where the pulse has reached the z-normal DFT planes and every PML
    slab holds nonzero field -- and linearity (source x2 gives fields x2
    exactly, since scaling by 2 commutes with rounding) after 60 steps.  Rounding is relative only
    above FLT_MIN: the wavefront's leading edge decays into subnormals, where the grid is absolute
    and x2 is no longer exact.  Those differences (measured: at most 4.7e-38 absolute) then travel
    with the field and flip the last bit of a few values up to ~1e-31 (measured: 199 values, 2.9e-7
    relative); far above that a flip needs a 1e-12-probability coincidence, so values of magnitude
    >= 1e-25 must match exactly."""
    Nt = 440
    print(f"[case] 400x400x200 parity, {Nt} steps", flush=True)
    model = workloads.dipole_slab(Nt=Nt, monitors=True)
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=16)
    p_g, m_g, f_g, info = _hip_session(model, np.float32, chunks=[263, Nt - 263])
    for k in f_c:
        np.testing.assert_array_equal(f_g[k], f_c[k], err_msg=k)
    for a, b in zip(p_g, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    _assert_monitors_equal(model, p_g, m_g, p_c, m_c)
    # what was compared is not trivially zero: every 12-cell PML slab and every DFT plane
    n = 12
    for k in ("ez", "hx"):
        v = f_g[k]
        for sl in ((slice(0, n),), (slice(-n, None),), (slice(None), slice(0, n)), (slice(None), slice(-n, None)),
                   (slice(None), slice(None), slice(0, n)), (slice(None), slice(None), slice(-n, None))):
            assert np.abs(v[sl]).max() > 0, (k, sl)
    for name, m in model.monitors.items():
        assert np.abs(m["values"]).max() > 0, name
    del f_c, f_g
    print("[case] 400x400x200 linearity", flush=True)
    model = workloads.dipole_slab(Nt=60, monitors=False)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    res = []
    for scale in (1.0, 2.0):
        pr = [dict(p, values=(p["values"] * np.float32(scale)).astype(np.float32)) for p in probes]
        with Session(c, Nx, Ny, Nz) as s:
            s.set_sources(pr, [], Nt)
            s.run(Nt)
            s.synchronize()
            res.append(s.fields(("ez", "hx")))
    for k in ("ez", "hx"):
        big = np.abs(res[0][k]) >= 1e-25
        np.testing.assert_array_equal(res[1][k][big], 2 * res[0][k][big])
        np.testing.assert_allclose(res[1][k][~big], 2 * res[0][k][~big], rtol=1e-6, atol=1e-36)
        assert rel_l2(res[1][k], 2 * res[0][k]) < 1e-7
    assert np.abs(res[0]["ez"]).max() > 0


@pytest.mark.parametrize("dtype,name", [(np.float32, "tiny_fp32.npz"), (np.float64, "tiny_fp64.npz")])
def test_hip_reproduces_golden(hip_ready, dtype, name):
    """The committed oracle fixtures, reproduced by the HIP engine bit for bit."""
    from pathlib import Path

    print(f"[case] golden {name}", flush=True)
    want = np.load(Path(__file__).resolve().parent / "golden" / name)
    model = workloads.tiny_model(Nt=60)
    p_g, m_g, f_g, _ = _hip_session(model, dtype)
    model.unpack(p_g, m_g, gpu=True)
    for i, p in enumerate(p_g):
        np.testing.assert_array_equal(p["values"], want[f"probe_{i}"])
    for k, m in model.monitors.items():
        np.testing.assert_array_equal(m["values"], want[f"monitor_{k}"], err_msg=k)
    for k, v in f_g.items():
        np.testing.assert_array_equal(v, want[f"field_{k}"], err_msg=k)


@pytest.mark.parametrize("variant", [0, 1, 4, 6, 7, 9, 10])
def test_boundary_plane_monitors(hip_ready, variant):
    """Time-domain monitors of all six components on the first and last planes each axis allows
    (36 monitors): identical to the oracle whether the sample is read from the stored totals (the
    separate kernels) or summed from the splits (the fused pass)."""
    print(f"[case] boundary monitors variant={variant}", flush=True)
    model = workloads.tiny_model(Nt=30, time_monitors=False, phasor_monitors=False)
    for f in ("ex", "ey", "ez", "hx", "hy", "hz"):
        for ai, axis in enumerate("xyz"):
            n = model.fshape[f][ai]
            for idx in (0, n - 2):
                model.add_field_monitor(f"b_{f}_{axis}{idx}", f, axis, index=idx, n_step=1)
    p_c, m_c, _ = run_oracle(model, np.float32, n_threads=2, fields=False)
    p_g, m_g, _, _ = _hip_session(model, np.float32, variant=variant, chunks=[17, 13])
    _assert_monitors_equal(model, p_g, m_g, p_c, m_c)


@pytest.mark.parametrize("dtype,flags", [(np.float32, 0), (np.float32, _lib.COEF_RAW), (np.float64, 0)])
def test_graded_grid_every_coefficient_kind(hip_ready, dtype, flags):
    """A graded grid holding every coefficient kind, through the session's default kernels and every
    fused, two-step and vector-column variant (Nz = 16): probes, monitors and fields bit-exact
    against the oracle."""
    print(f"[case] graded grid {np.dtype(dtype).name} flags={flags}", flush=True)
    model = _graded_everything(Nt=70, coef_dtype=dtype)
    p_c, m_c, f_c = run_oracle(model, dtype, n_threads=3)
    for variant in (None, 2, 3, 5, 6, 7, 8, 9, 13, 14, 15, 10, 11, 12):
        p_g, m_g, f_g, _ = _hip_session(model, dtype, flags, chunks=[41, 29], variant=variant)
        for a, b in zip(p_g, p_c):
            np.testing.assert_array_equal(a["values"], b["values"], err_msg=f"variant {variant}")
        _assert_monitors_equal(model, p_g, m_g, p_c, m_c)
        for k in f_c:
            np.testing.assert_array_equal(f_g[k], f_c[k], err_msg=f"{k} variant {variant}")
