"""Touchstone files: the reader follows rfnetwork/touchstone.py (S21 before S12 in two-port files,
DB/MA/RI formats, frequency units, a trailing noise block), checked on a measured amplifier file
from the reference's own test data; the writer round-trips through it for 1..5 ports and every
format."""
from pathlib import Path

import numpy as np
import pytest

from rfnetwork_amd import touchstone

FIXTURE = Path(__file__).parent / "golden" / "QPL9503_5V_with_NP.s2p"  # rfnetwork tests/data


def _text_rows():
    rows = []
    for line in FIXTURE.read_text().splitlines():
        line = line.strip()
        if line and line[0] not in "!#":
            rows.append([float(v) for v in line.split()])
    return rows


def test_reads_measured_amplifier_file():
    S, noise, comments = touchstone.read_snp(FIXTURE)
    rows = _text_rows()
    s_rows = [r for r in rows if len(r) == 9]
    n_rows = [r for r in rows if len(r) == 5]
    assert S.shape == (len(s_rows), 2, 2)
    np.testing.assert_array_equal(S.coords["frequency"], [r[0] for r in s_rows])  # the file is in Hz
    r = s_rows[0]
    db_ang = lambda d, a: 10 ** (d / 20) * np.exp(1j * np.deg2rad(a))
    # pairs in the file: S11, S21, S12, S22; the array is [f, b, a]
    assert S[0, 0, 0] == pytest.approx(db_ang(r[1], r[2]), rel=1e-12)
    assert S[0, 1, 0] == pytest.approx(db_ang(r[3], r[4]), rel=1e-12)  # S21: the amplifier's gain
    assert S[0, 0, 1] == pytest.approx(db_ang(r[5], r[6]), rel=1e-12)
    assert S[0, 1, 1] == pytest.approx(db_ang(r[7], r[8]), rel=1e-12)
    assert 20 * np.log10(abs(S[0, 1, 0])) > 20  # gain, not isolation
    assert noise.shape == (len(n_rows), 3)
    n0 = n_rows[0]
    assert noise.coords["frequency"][0] == n0[0]
    assert noise[0, 0].real == pytest.approx(10 ** (n0[1] / 10))
    assert noise[0, 1] == pytest.approx(n0[2] * np.exp(1j * np.deg2rad(n0[3])))
    assert noise[0, 2].real == pytest.approx(n0[4])
    assert comments[0] == "QPL9503"


@pytest.mark.parametrize("N", [1, 2, 3, 5])
@pytest.mark.parametrize("fmt", ["ri", "ma", "db"])
@pytest.mark.parametrize("unit", ["hz", "ghz"])
def test_write_read_round_trip(tmp_path, N, fmt, unit):
    rng = np.random.default_rng(N * 7 + len(fmt))
    f = np.linspace(1e9, 1e10, 11)
    S = rng.standard_normal((11, N, N)) + 1j * rng.standard_normal((11, N, N))
    path = tmp_path / f"x.s{N}p"
    touchstone.write_snp(path, S, f, fmt=fmt, freq_unit=unit, comments=["written by rfnetwork_amd"])
    S2, noise, comments = touchstone.read_snp(path)
    assert noise is None and comments == ["written by rfnetwork_amd"]
    np.testing.assert_allclose(S2.coords["frequency"], f, rtol=1e-11)
    np.testing.assert_allclose(np.asarray(S2), S, rtol=1e-9, atol=1e-12)


def test_two_port_order_in_file(tmp_path):
    S = np.array([[[1.0, 2.0], [3.0, 4.0]]])  # S11 = 1, S12 = 2, S21 = 3, S22 = 4
    path = tmp_path / "t.s2p"
    touchstone.write_snp(path, S, [1e9])
    data = [ln for ln in path.read_text().splitlines() if ln and ln[0] not in "!#"][0].split()
    assert [float(v) for v in data[1::2]] == [1.0, 3.0, 2.0, 4.0]


def test_rejects_mismatched_extension(tmp_path):
    with pytest.raises(ValueError):
        touchstone.write_snp(tmp_path / "t.s3p", np.zeros((1, 2, 2)), [1e9])
    with pytest.raises(ValueError):
        touchstone.read_snp(tmp_path / "t.txt")
