"""The coupled-stripline closed form against what the mathematics fixes: far-apart strips are two
independent striplines, each with Wheeler's wide-strip impedance 30 pi b / (sqrt(er) (w + 0.441 b))
(the fringing term 2 b ln 2 / pi = 0.441 b); the odd mode lies below the even mode; both scale as
1/sqrt(er); a strip of zero thickness gives the same result with t = 0 passed or not."""
import numpy as np

from rfnetwork_amd.utils import coupled_sline_impedance, ustrip_impedance


def test_far_apart_strips_are_single_striplines():
    w, b, er = 0.02, 0.06, 3.66
    zo, ze = coupled_sline_impedance(w, 50 * b, b, er)
    wheeler = 30 * np.pi * b / (np.sqrt(er) * (w + 2 * np.log(2) / np.pi * b))
    assert abs(zo - wheeler) / wheeler < 1e-9 and abs(ze - wheeler) / wheeler < 1e-9


def test_mode_order_and_permittivity_scaling():
    for w, s in ((0.02, 0.01), (0.03, 0.005), (0.04, 0.01)):
        zo, ze = coupled_sline_impedance(w, s, 0.06, 3.66)
        assert zo < ze
        zo1, ze1 = coupled_sline_impedance(w, s, 0.06, 1.0)
        np.testing.assert_allclose([zo1 / zo, ze1 / ze], np.sqrt(3.66), rtol=1e-12)
        assert coupled_sline_impedance(w, s, 0.06, 3.66, t=0) == (zo, ze)


def test_microstrip_static_limit_is_the_air_line_and_wheeler_hammerstad():
    """Thin strip, low frequency: the impedance in air is eta0/(2 pi) ln(F/u + sqrt(1 + 4/u^2)) with
    F -> 6 for narrow strips (u = 1: 126.5 ohm, Wheeler's tabulated 126 ohm); on a substrate it lies
    within 1% of the Hammerstad 1975 forms (eps_eff = (er+1)/2 + (er-1)/2 / sqrt(1 + 12/u),
    Z0 = eta0 / (sqrt(eps_eff) (u + 1.393 + 0.667 ln(u + 1.444))) for u >= 1, the 60 ln(8/u + u/4)
    form below)."""
    z_air, e_air = ustrip_impedance(0.02, 0.02, 1.0 + 1e-9, 1e3, t=1e-12)
    assert abs(e_air - 1) < 1e-8 and abs(z_air - 126.5) < 0.2
    for u in (0.3, 0.7, 1.0, 2.0, 5.0):
        for er in (2.2, 3.66, 10.2):
            z, e = ustrip_impedance(u * 0.02, 0.02, er, 1e3, t=1e-12)
            e75 = (er + 1) / 2 + (er - 1) / 2 / np.sqrt(1 + 12 / u)
            z75 = (376.73 / (np.sqrt(e75) * (u + 1.393 + 0.667 * np.log(u + 1.444))) if u >= 1
                   else 60 / np.sqrt(e75) * np.log(8 / u + u / 4))
            assert abs(e - e75) / e75 < 0.01 and abs(z - z75) / z75 < 0.01, (u, er, z, z75, e, e75)


def test_microstrip_dispersion_and_thickness_trends():
    """eps_eff rises with frequency from its static value towards er; copper thickness widens the
    strip electrically and lowers Z0; at 10 GHz the 0.04 inch strip on 0.02 inch of er 3.66 (the solver
    test's line) sits at 52.72 ohm."""
    f = np.array([1e3, 1e9, 5e9, 10e9, 20e9, 40e9, 400e9])
    z, e = ustrip_impedance(0.04, 0.02, 3.66, f)
    assert np.all(np.diff(e) > 0) and e[-1] < 3.66 and abs(e[-1] - 3.66) < 0.1
    assert ustrip_impedance(0.04, 0.02, 3.66, 1e3, t=0.0026)[0] < ustrip_impedance(0.04, 0.02, 3.66, 1e3, t=0.0013)[0]
    assert abs(ustrip_impedance(0.04, 0.02, 3.66, 10e9)[0] - 52.716) < 0.01
