"""Design-region coefficient assembly on the GPU (rfdtd_set_design_permittivity): a session built
from one permittivity of the design box and updated on the device to another runs, bit for bit,
as a session built by the host assembly from the second permittivity (rfnetwork/solver.py:892-961
averaging and Ca/Cb, :1463-1523 packing)."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

pytestmark = pytest.mark.gpu

DESIGN = (8, 8, 6)


def _model(eps, coef_dtype=np.float32, Nt=120, **kw):
    return workloads.design_strip(eps, Nt=Nt, nx=48, ny=32, nz=20, design=DESIGN, n_pml=5, h=3,
                                  coef_dtype=coef_dtype, pulse=(20e9, 80e-12, 60e-12), **kw)


def _eps(seed):
    return np.random.default_rng(seed).uniform(1.0, 4.0, size=DESIGN)


def _run(session, model, Nt):
    _, probes, _, _, *_ = model.pack(gpu=True, dtype=session.dtype, coefficients=False)
    session.set_sources(probes, [], Nt)
    session.run(Nt)
    session.synchronize()
    session.read()
    return [np.array(p["values"]) for p in probes], session.fields()


def _reference(model, dtype, flags, Nt):
    coeff, _, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True, dtype=dtype)
    with Session(coeff, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        return _run(s, model, Nt)


def _assert_same(a, b):
    for x, y in zip(a[0], b[0]):
        np.testing.assert_array_equal(x, y)
    for k in a[1]:
        np.testing.assert_array_equal(a[1][k], b[1][k], err_msg=k)


@pytest.mark.parametrize("dtype,flags", [(np.float32, 0), (np.float32, _lib.COEF_RAW), (np.float64, 0)])
def test_design_update_equals_host_assembly(hip_ready, dtype, flags):
    """One session updated twice (the second update rewrites the box's entries in place) against
    sessions built from the host assembly of each permittivity."""
    print(f"[case] design update dtype={np.dtype(dtype).name} flags={flags}", flush=True)
    Nt = 120
    m0, m1, m2 = (_model(_eps(s), coef_dtype=dtype, Nt=Nt) for s in (1, 2, 3))
    cbox = workloads.design_box(48, 32, DESIGN)
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True, dtype=dtype)
    for target, eps in ((m1, _eps(2)), (m2, _eps(3))):
        with Session(coeff, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
            s.set_design_permittivity(m0, cbox, _eps(2))
            if target is m2:
                s.set_design_permittivity(m0, cbox, eps)
            got = _run(s, m0, Nt)
        _assert_same(got, _reference(target, dtype, flags, Nt))


def test_design_update_two_boxes(hip_ready):
    """Two different boxes one after the other: the cells of the first keep their entries."""
    print("[case] design two boxes", flush=True)
    Nt = 100
    cbox = workloads.design_box(48, 32, DESIGN)
    x0, x1, y0, y1, z0, z1 = cbox
    e1, e2 = _eps(4), _eps(5)
    m0 = _model(_eps(1), Nt=Nt)
    # the expected model: the whole box from e1, then its lower half in x from e2
    expect = e1.copy()
    expect[:4] = e2[:4]
    m_exp = _model(expect, Nt=Nt)
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True)
    with Session(coeff, Nx, Ny, Nz) as s:
        s.set_design_permittivity(m0, cbox, e1)
        s.set_design_permittivity(m0, (x0, x0 + 4, y0, y1, z0, z1), e2[:4])
        got = _run(s, m0, Nt)
    _assert_same(got, _reference(m_exp, np.float32, 0, Nt))


def test_design_update_in_a_batch(hip_ready):
    """Updating one member of a batch changes that member only."""
    print("[case] design batch", flush=True)
    Nt = 100
    cbox = workloads.design_box(48, 32, DESIGN)
    ma, mb, mb2 = _model(_eps(6), Nt=Nt), _model(_eps(7), Nt=Nt), _model(_eps(8), Nt=Nt)
    grids = []
    for m in (ma, mb):
        coeff, _, _, _, Nx, Ny, Nz, _ = m.pack(gpu=True)
        grids.append((coeff, Nx, Ny, Nz, 1))
    with Batch(grids) as b:
        b.set_design_permittivity(1, mb, cbox, _eps(8))
        for k, m in enumerate((ma, mb)):
            _, probes, _, _, *_ = m.pack(gpu=True, coefficients=False)
            b.set_sources(k, probes, [], Nt)
        b.run(Nt)
        b.synchronize()
        b.read()
        fields = [b.fields(k) for k in range(2)]
    for m, f in ((ma, fields[0]), (mb2, fields[1])):
        ref = _reference(m, np.float32, 0, Nt)[1]
        for k in ref:
            np.testing.assert_array_equal(f[k], ref[k], err_msg=k)


def test_design_update_refuses_non_plain_edges(hip_ready):
    """A box reaching a conductor, port or PML edge is refused and the session is left unchanged;
    so is a stated current permittivity that the session does not hold."""
    print("[case] design refusal", flush=True)
    Nt = 60
    m0 = _model(_eps(1), Nt=Nt)
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True)
    with Session(coeff, Nx, Ny, Nz) as s:
        with pytest.raises(Exception, match="plain dielectric"):
            s.set_design_permittivity(m0, (1, 4, 1, 4, 0, 3), np.full((3, 3, 3), 2.0))  # in the PML
        x0, x1, y0, y1, z0, z1 = workloads.design_box(48, 32, DESIGN)
        with pytest.raises(Exception, match="plain dielectric"):  # reaches the feed conductor at x0 - 1
            s.set_design_permittivity(m0, (x0 - 2, x0, y0, y1, z0, z1), np.full((2, 8, 6), 2.0))
        got = _run(s, m0, Nt)
    _assert_same(got, _reference(m0, np.float32, 0, Nt))


@pytest.mark.parametrize("mode", ["history", "checkpoint"])
def test_design_loop_equals_rebuilt_models(hip_ready, mode):
    """Successive evaluations on one resident session give J, dJ/deps_r and S equal, bit for bit,
    to eps_r_gradient on models rebuilt with each permittivity."""
    print(f"[case] design loop {mode}", flush=True)
    from rfnetwork_amd import adjoint
    from rfnetwork_amd.design import DesignLoop

    Nt = 300
    freq = (16e9, 20e9, 24e9)
    cbox = workloads.design_box(48, 32, DESIGN)
    m0 = _model(_eps(1), Nt=Nt, frequency=freq)
    with DesignLoop(m0, cbox, freq, mode=mode, checkpoint_every=70) as loop:
        for seed in (2, 3):
            J, g, S = loop.evaluate(_eps(seed))
            mr = _model(_eps(seed), Nt=Nt, frequency=freq)
            Jr, gr, Sr = adjoint.eps_r_gradient(mr, cbox, freq, mode=mode, checkpoint_every=70)
            assert J == Jr
            np.testing.assert_array_equal(g, gr)
            np.testing.assert_array_equal(S, Sr)
            assert loop.timing["update_ms"] > 0


def _graded_model(eps_box, Nt):
    """A graded grid (every cell width different), lossy random dielectric, PML on two sides and a
    port outside a 6x5x4 design box at cells [8, 14) x [6, 11) x [5, 9)."""
    from rfnetwork_amd.geometry import Rectangle
    from rfnetwork_amd.solver import FDTD_Solver
    from rfnetwork_amd.geometry import Box

    rng = np.random.default_rng(21)
    d = 1e-4 * 39.3701
    widths = [d * rng.uniform(0.7, 1.5, size=n) for n in (24, 18, 14)]
    edges = [np.concatenate([[0.0], np.cumsum(w)]) for w in widths]
    s = FDTD_Solver(Box((0, 0, 0), tuple(float(e[-1]) for e in edges)))
    s.set_grid(*edges)
    vol = rng.uniform(1.0, 4.0, size=(24, 18, 14))
    vol[8:14, 6:11, 5:9] = eps_box
    s.add_permittivity_volume(vol, (0, 0, 0), sigma=rng.uniform(0.0, 0.05, size=(24, 18, 14)))
    ex, ey, ez = edges
    s.add_lumped_port(1, Rectangle([(ex[4], ey[4], ez[2]), (ex[4], ey[7], ez[2]), (ex[4], ey[7], ez[5])]), "z+")
    s.assign_PML_boundaries("x-", "x+", n_pml=3)
    s.generate_mesh(courant=0.9)
    src = s.gaussian_modulated_source(f0=20e9, width=60e-12, t0=30e-12, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(1e-2 * np.asarray(src), 1)
    return s


def test_design_update_on_a_graded_grid(hip_ready):
    """Cell widths that differ from cell to cell: the device averaging forms the weights' sum in
    float32 as the host assembly does, so the updated session runs as the rebuilt model's."""
    print("[case] design graded grid", flush=True)
    Nt = 80
    rng = np.random.default_rng(5)
    e0, e1 = rng.uniform(1.0, 4.0, size=(2, 6, 5, 4))
    m0, m1 = _graded_model(e0, Nt), _graded_model(e1, Nt)
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True)
    with Session(coeff, Nx, Ny, Nz) as s:
        s.set_design_permittivity(m0, (8, 14, 6, 11, 5, 9), e1)
        got = _run(s, m0, Nt)
    _assert_same(got, _reference(m1, np.float32, 0, Nt))


def _device(a, dtype):
    import torch

    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("dtype,flags,tdt", [(np.float32, 0, "float32"), (np.float32, _lib.COEF_RAW, "float64"),
                                             (np.float64, 0, "float64")])
def test_design_update_from_device_memory(hip_ready, dtype, flags, tdt):
    """rfdtd_design_prepare + rfdtd_set_design_permittivity_device, the permittivity read from a torch
    tensor in device memory (twice: the second update rewrites the box in place), against the host
    update with the same values: identical probes and fields."""
    import torch

    print(f"[case] design update from device memory dtype={np.dtype(dtype).name} flags={flags} tensor={tdt}",
          flush=True)
    Nt = 120
    tdtype = getattr(torch, tdt)
    # values a float32 tensor holds exactly, so that the host update (float64) sees the same numbers
    e2, e3 = (_eps(s).astype(np.float32).astype(np.float64) for s in (2, 3))
    m0 = _model(_eps(1), coef_dtype=dtype, Nt=Nt)
    cbox = workloads.design_box(48, 32, DESIGN)
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True, dtype=dtype)
    with Session(coeff, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        s.set_design_permittivity(m0, cbox, e2)
        s.set_design_permittivity(m0, cbox, e3)
        want = _run(s, m0, Nt)
    with Session(coeff, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        s.set_design_permittivity(m0, cbox, _device(e2, tdtype))
        t3 = _device(e3, tdtype)
        s.set_design_permittivity(m0, cbox, t3)
        got = _run(s, m0, Nt)
    _assert_same(got, want)


def test_design_update_on_caller_streams(hip_ready):
    """Steps on one caller stream, the update from device memory on another, the next steps on the
    first (both streams non-blocking): the update waits for the earlier run and the next run waits
    for the update, so probes and fields equal those of a host update between the same steps."""
    import ctypes

    import torch

    from rfnetwork_amd import design

    print("[case] design update on caller streams", flush=True)
    Nt, n1 = 120, 47
    e2 = _eps(2).astype(np.float32).astype(np.float64)
    m0 = _model(_eps(1), Nt=Nt)
    cbox = workloads.design_box(48, 32, DESIGN)
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True)

    def steps(s, run_a, update, run_b):
        _, probes, _, _, *_ = m0.pack(gpu=True, coefficients=False)
        s.set_sources(probes, [], Nt)
        run_a()
        update()
        run_b()
        s.synchronize()
        s.read()
        return [np.array(p["values"]) for p in probes], s.fields()

    with Session(coeff, Nx, Ny, Nz) as s:
        want = steps(s, lambda: (s.run(n1), s.synchronize()), lambda: s.set_design_permittivity(m0, cbox, e2),
                     lambda: s.run(Nt - n1))
    hip = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime the library is bound to
    streams = []
    for _ in range(2):
        h = ctypes.c_void_p()
        assert hip.hipStreamCreateWithFlags(ctypes.byref(h), ctypes.c_uint(1)) == 0  # non-blocking
        streams.append(h)
    a, b = streams
    t = torch.tensor(e2, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    try:
        with Session(coeff, Nx, Ny, Nz) as s:
            design.prepare_design_box(s._h, s._L, 0, m0, cbox, {}, np.float32)
            box = (ctypes.c_int32 * 6)(*cbox)

            def update():
                _lib.check(s._L.rfdtd_set_design_permittivity_device(
                    s._h, 0, ctypes.cast(box, ctypes.c_void_p), ctypes.c_void_p(t.data_ptr()), 0, b),
                    "rfdtd_set_design_permittivity_device")

            got = steps(s, lambda: s.run(n1, stream=a.value), update, lambda: s.run(Nt - n1, stream=a.value))
    finally:
        for h in streams:
            hip.hipStreamDestroy(h)
    _assert_same(got, want)


def test_design_update_from_device_memory_refusals(hip_ready):
    """A device update needs a prepared box and cells inside it; a box on a conductor edge is refused
    when it is prepared."""
    import ctypes

    print("[case] design device refusals", flush=True)
    from rfnetwork_amd import design

    m0 = _model(_eps(1), Nt=60)
    cbox = workloads.design_box(48, 32, DESIGN)
    x0, x1, y0, y1, z0, z1 = cbox
    coeff, _, _, _, Nx, Ny, Nz, _ = m0.pack(gpu=True)
    with Session(coeff, Nx, Ny, Nz) as s:
        import torch

        t = torch.ones(DESIGN, dtype=torch.float64, device="cuda") * 2.0
        with pytest.raises(ValueError, match="no design box prepared"):
            design.set_design_permittivity_device(s._h, s._L, 0, cbox, t)
        with pytest.raises(ValueError, match="plain dielectric"):
            design.prepare_design_box(s._h, s._L, 0, m0, (x0 - 2, x0, y0, y1, z0, z1), {}, np.float32)
        design.prepare_design_box(s._h, s._L, 0, m0, cbox, {}, np.float32)
        far = (x0 + 3, x1 + 3, y0, y1, z0, z1)
        b = (ctypes.c_int32 * 6)(*far)
        rc = s._L.rfdtd_set_design_permittivity_device(s._h, 0, ctypes.cast(b, ctypes.c_void_p),
                                                       ctypes.c_void_p(t.data_ptr()), 1, None)
        assert rc == 1 and b"outside the prepared" in s._L.rfdtd_last_error()


def test_design_loop_device_tensors(hip_ready):
    """A DesignLoop fed torch tensors in device memory (float32 and float64) gives J, S and dJ/deps_r
    identical to the same loop fed numpy arrays; the autograd Function on it hands torch that
    gradient on the tensor's device."""
    import torch

    from rfnetwork_amd.adjoint import EpsRObjective
    from rfnetwork_amd.design import DesignLoop

    print("[case] design loop from device tensors", flush=True)
    Nt = 300
    freq = (16e9, 20e9, 24e9)
    cbox = workloads.design_box(48, 32, DESIGN)
    vals = [_eps(s).astype(np.float32).astype(np.float64) for s in (2, 3, 4)]
    m0 = _model(_eps(1), Nt=Nt, frequency=freq)
    with DesignLoop(m0, cbox, freq, mode="history") as host_loop:
        want = [host_loop.evaluate(v) for v in vals]
    m0 = _model(_eps(1), Nt=Nt, frequency=freq)
    with DesignLoop(m0, cbox, freq, mode="history") as loop:
        for v, tdt, (Jw, gw, Sw) in zip(vals, (torch.float32, torch.float64, torch.float32), want):
            J, g, S = loop.evaluate(_device(v, tdt))
            assert J == Jw
            np.testing.assert_array_equal(g, gw)
            np.testing.assert_array_equal(S, Sw)
        # a host array after device updates, then the autograd Function with a device tensor
        J, g, _ = loop.evaluate(vals[0])
        assert J == want[0][0]
        np.testing.assert_array_equal(g, want[0][1])
        t = _device(vals[1], torch.float64).requires_grad_(True)
        Jt = EpsRObjective.apply(t, loop)
        Jt.backward()
        assert float(Jt.detach()) == want[1][0] and t.grad.is_cuda
        np.testing.assert_array_equal(t.grad.cpu().numpy(), want[1][1])
