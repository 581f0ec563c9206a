"""Writes the golden fixtures of the FDTD hot path from the CPU oracle.

tiny_fp32.npz / tiny_fp64.npz: the 14x11x9 model of rfnetwork_amd.workloads.tiny_model (PML on all
faces, random dielectric block, PEC strip, two ports, a voltage probe, 4 time-domain and 3 running-DFT
monitors), 60 steps: every probe series, every monitor padded to full Yee planes, and all 18 fields.
The reference ships no value-level fixtures for its solver; these pin the oracle (tests check it still
produces them) and are the bit-exact target of the HIP engine on the GPU.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from helpers import run_oracle  # noqa: E402
from rfnetwork_amd import workloads  # noqa: E402


def golden_arrays(dtype):
    model = workloads.tiny_model(Nt=60)
    probes, monitors, fields = run_oracle(model, dtype, n_threads=3)
    model.unpack(probes, monitors, gpu=False)
    out = {f"probe_{i}": p["values"] for i, p in enumerate(probes)}
    out.update({f"monitor_{k}": m["values"] for k, m in model.monitors.items()})
    out.update({f"field_{k}": v for k, v in fields.items()})
    return out


def main():
    for dtype, name in ((np.float32, "tiny_fp32.npz"), (np.float64, "tiny_fp64.npz")):
        np.savez_compressed(HERE / name, **golden_arrays(dtype))
        print("wrote", HERE / name)


if __name__ == "__main__":
    main()
