"""The C-ABI library loads and exports every entry point declared in include/rfdtd.h; the oracle
library exports its entry points.  No compute call is made (no GPU needed)."""
import ctypes
import re
import subprocess
from pathlib import Path

from rfnetwork_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rfdtd_[a-z_0-9]+)\s*\(", text)))


def _exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_library_exports_every_declared_symbol():
    declared = _declared("rfdtd.h")
    assert "rfdtd_solver_run" in declared and len(declared) >= 15
    exported = _exported(_lib.LIB_PATH)
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared)


def test_library_loads_and_reports_version():
    L = _lib.lib()
    assert L.rfdtd_version() == 3
    assert L.rfdtd_last_error() == b""


def test_oracle_exports():
    exported = _exported(ROOT / "oracle" / "liboracle_fdtd.so")
    for s in ("oracle_run_f32", "oracle_run_f64", "oracle_last_error"):
        assert s in exported


def test_abi_struct_layout():
    assert ctypes.sizeof(_lib.Probe) == 32
    assert ctypes.sizeof(_lib.Monitor) == 40
