"""Touchstone (.sNp) files: reading, as rfnetwork/touchstone.py:7-155 does, and writing the
S-parameters that the solver produces, so that solved models can be handed to the circuit side.

Format (Touchstone 1.x): an option line ``# <freq unit> S <RI|MA|DB> R <ref>``, ``!`` comments, then
per frequency the frequency followed by N^2 value pairs.  Two-port files list S11 S21 S12 S22;
larger files list the matrix row by row (S11 S12 ... S1N, S21 ...), each row starting on a new
line with at most four pairs per line.  Noise parameters may follow a two-port block: five values
per frequency (frequency, NFmin dB, |Gopt|, angle(Gopt) degrees, Rn normalised), recognised by the
frequency restarting lower.
"""
import re
from pathlib import Path

import numpy as np

from .utils import LabeledArray

_FREQ_UNIT = {"hz": 1e0, "khz": 1e3, "mhz": 1e6, "ghz": 1e9}


def _n_ports(path):
    m = re.fullmatch(r"\.s(\d+)p", Path(path).suffix, re.IGNORECASE)
    if m is None:
        raise ValueError("Invalid extension for touchstone file. Expected *.sNp, where N is the number of ports.")
    return int(m.group(1))


def read_snp(filepath):
    """Returns (S LabeledArray [frequency, b, a], noise LabeledArray [frequency, noise_param] or None,
    comments).  Noise parameters: nf_min (linear), gamma_opt (complex), rn (normalised)."""
    N = _n_ports(filepath)
    comments, tokens, options = [], [], []
    with open(filepath, "r", encoding="utf-8", errors="ignore") as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line[0] == "#":
                if not options:  # a second option line (some noise blocks) is ignored
                    options = [x.lower() for x in line[1:].split()]
            elif line[0] == "!":
                comments.append(line[1:])
            else:
                tokens += line.split("!")[0].split()
    # defaults of the format: GHz, S, MA, R 50
    unit = options[0] if len(options) > 0 else "ghz"
    fmt = options[2] if len(options) > 2 else "ma"
    values = np.asarray(tokens, dtype=np.float64)
    stride = 2 * N * N + 1
    freq = values[::stride]
    drop = np.nonzero(np.diff(freq) < 0)[0]
    noise = None
    if len(drop):
        n_s = int(drop[0]) + 1
        s_vals, n_vals = values[: n_s * stride], values[n_s * stride:]
        rows = n_vals.reshape(-1, 5)
        nf_min = 10 ** (rows[:, 1] / 10)
        gamma_opt = rows[:, 2] * np.exp(1j * np.deg2rad(rows[:, 3]))
        noise = LabeledArray(np.column_stack([nf_min, gamma_opt, rows[:, 4]]),
                             coords=dict(frequency=rows[:, 0] * _FREQ_UNIT[unit],
                                         noise_param=["nf_min", "gamma_opt", "rn"]))
    else:
        s_vals = values
    s = s_vals.reshape(-1, stride)
    f_hz = s[:, 0] * _FREQ_UNIT[unit]
    pairs = s[:, 1:].reshape(-1, N, N, 2)
    p1, p2 = pairs[..., 0], pairs[..., 1]
    if fmt == "ri":
        S = p1 + 1j * p2
    elif fmt == "ma":
        S = p1 * np.exp(1j * np.deg2rad(p2))
    else:  # db
        S = 10 ** (p1 / 20) * np.exp(1j * np.deg2rad(p2))
    if N == 2:  # S21 precedes S12 in two-port files
        S = np.transpose(S, (0, 2, 1))
    S = LabeledArray(S, coords=dict(frequency=f_hz, b=np.arange(1, N + 1), a=np.arange(1, N + 1)))
    return S, noise, comments


def write_snp(filepath, S, frequency=None, fmt="ri", ref_z=50.0, freq_unit="hz", comments=()):
    """Write S [frequency, b, a] (N x N per frequency) to an .sNp file whose N matches.  frequency
    in Hz defaults to S.coords["frequency"]."""
    S = np.asarray(S)
    if S.ndim != 3 or S.shape[1] != S.shape[2]:
        raise ValueError("S must have shape [frequency, N, N]")
    N = S.shape[1]
    if _n_ports(filepath) != N:
        raise ValueError(f"file extension does not match the {N}-port data")
    if frequency is None:
        frequency = getattr(S, "coords", {}).get("frequency")
        if frequency is None:
            raise ValueError("frequency is required")
    frequency = np.asarray(frequency, dtype=np.float64)
    fmt, unit = fmt.lower(), freq_unit.lower()
    if fmt not in ("ri", "ma", "db") or unit not in _FREQ_UNIT:
        raise ValueError("fmt must be ri, ma or db and freq_unit hz, khz, mhz or ghz")
    if fmt == "ri":
        p1, p2 = S.real, S.imag
    elif fmt == "ma":
        p1, p2 = np.abs(S), np.angle(S, deg=True)
    else:
        p1, p2 = 20 * np.log10(np.abs(S)), np.angle(S, deg=True)
    lines = [f"!{c}" for c in comments]
    lines.append(f"# {unit.upper()} S {fmt.upper()} R {ref_z:g}")
    for k, f in enumerate(frequency):
        if N == 2:
            order = [(0, 0), (1, 0), (0, 1), (1, 1)]
            body = " ".join(f"{p1[k, b, a]:.12g} {p2[k, b, a]:.12g}" for b, a in order)
            lines.append(f"{f / _FREQ_UNIT[unit]:.12g} {body}")
            continue
        first = True
        for b in range(N):
            row = [f"{p1[k, b, a]:.12g} {p2[k, b, a]:.12g}" for a in range(N)]
            for j in range(0, N, 4):
                chunk = " ".join(row[j:j + 4])
                lines.append(f"{f / _FREQ_UNIT[unit]:.12g} {chunk}" if first else chunk)
                first = False
    Path(filepath).write_text("\n".join(lines) + "\n")
