"""The host side of the on-device coefficient assembly: a model meshed with
coefficients="device" (no full-grid arrays on the host) has the same time step, cell arrays, port
source gains, lumped-element terms and edge-correction factors as the model meshed with the host
assembly, so the description handed to the GPU is the same; and the edge permittivity formed for a
few edges alone (the ports') equals the host assembly's full arrays (rfnetwork/solver.py:892-912)."""
import numpy as np
import pytest

from rfnetwork_amd import assembly, workloads

from helpers import graded_everything


def _pair(name, dtype):
    if name == "tiny":
        return [workloads.tiny_model(Nt=20, coef_dtype=dtype, coefficients=c) for c in ("host", "device")]
    if name == "microstrip":
        return [workloads.microstrip(Nt=20, coef_dtype=dtype, coefficients=c) for c in ("host", "device")]
    return [graded_everything(20, coef_dtype=dtype, coefficients=c) for c in ("host", "device")]


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", ["tiny", "microstrip", "graded"])
def test_device_mode_model_matches_host_mode(name, dtype):
    h, d = _pair(name, dtype)
    assert h.dt == d.dt and np.asarray(h.dt).dtype == np.asarray(d.dt).dtype
    np.testing.assert_array_equal(h.eps, d.eps)
    np.testing.assert_array_equal(h.sigma, d.sigma)
    for ph, pd in zip(h.ports + h.ports_inv, d.ports + d.ports_inv):
        assert (ph is None) == (pd is None)
        if ph is None:
            continue
        assert ph["idx"] == pd["idx"] and ph["field"] == pd["field"]
        assert np.asarray(ph["Vs_a"]).dtype == np.asarray(pd["Vs_a"]).dtype
        np.testing.assert_array_equal(_bits(ph["Vs_a"]), _bits(pd["Vs_a"]))
    assert list(h.probes) == list(d.probes)
    for (ah, eh, qh), (ad, ed, qd) in zip(h._lumped_fix, d._lumped_fix):
        assert ah == ad and eh == ed
        np.testing.assert_array_equal(_bits(qh), _bits(qd))
    assert len(h._edge_ops) == len(d._edge_ops)
    for (nh, ih, fh), (nd, id_, fd) in zip(h._edge_ops, d._edge_ops):
        assert nh == nd and ih == id_ and fh == fd and type(fh) is type(fd)
    # the descriptions handed to the GPU
    Dh, Dd = assembly.Description(h, dtype), assembly.Description(d, dtype)
    assert (Dh.struct.n_pml, Dh.struct.n_fix, Dh.n_fix_edges, Dh.n_scales) == \
        (Dd.struct.n_pml, Dd.struct.n_fix, Dd.n_fix_edges, Dd.n_scales)
    with pytest.raises(ValueError, match="device"):
        d.coefficient_arrays(gpu=True, dtype=dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_edge_eps_equals_host_arrays(dtype):
    model = graded_everything(20, coef_dtype=dtype)
    for f, full in (("ex", model.eps_ex), ("ey", model.eps_ey), ("ez", model.eps_ez)):
        got = model._edge_eps(f, (slice(None), slice(None), slice(None)))
        assert got.dtype == full.dtype
        np.testing.assert_array_equal(_bits(got), _bits(full), err_msg=f)
        blk = (slice(3, 9), 4, slice(2, 7))
        np.testing.assert_array_equal(_bits(model._edge_eps(f, blk)).ravel(), _bits(full[blk]).ravel())


def test_description_checks():
    """The description refuses a model whose coefficient dtype is not the session's (the arrays
    would be formed in the wrong type), and Batch.assembled needs one excitation count per model;
    both before any GPU call."""
    from rfnetwork_amd.batch import Batch

    m = workloads.tiny_model(Nt=10)
    with pytest.raises(ValueError, match="coef_dtype"):
        assembly.Description(m, np.float64)
    with pytest.raises(ValueError, match="one excitation count per model"):
        Batch.assembled([m, m], excitations=[1])
    D = assembly.Description(m, np.float32)
    assert D.struct.Nx == m.Nx and D.struct.n_pml == 6 and D.struct.dt == float(m.dt)
