"""The coupled-stripline closed form against what the mathematics fixes: far-apart strips are two
independent striplines, each with Wheeler's wide-strip impedance 30 pi b / (sqrt(er) (w + 0.441 b))
(the fringing term 2 b ln 2 / pi = 0.441 b); the odd mode lies below the even mode; both scale as
1/sqrt(er); a strip of zero thickness gives the same result with t = 0 passed or not."""
import numpy as np

from rfnetwork_amd.utils import coupled_sline_impedance


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
