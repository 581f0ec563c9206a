"""Launch-shape autotuning with a per-shape cache.

The engine's tuner (rfdtd_autotune) times every candidate kernel variant, block size and number of
x planes per workgroup on the session's own grids and keeps the fastest.  This module keys the
result by what determines it -- the grids (Nx, Ny, Nz, excitations), the element type and the
coefficient storage -- and keeps it in memory and in a JSON file, so a shape is tuned once per
machine.  The file is ``$RFNETWORK_AMD_TUNE_CACHE`` or ``~/.cache/rfnetwork_amd/launch.json``;
failing to read or write it only costs a re-tune.
"""
import ctypes
import hashlib
import json
import os
import pathlib

from . import _lib
from ._lib import check

_memory = {}


def cache_path():
    p = os.environ.get("RFNETWORK_AMD_TUNE_CACHE")
    return pathlib.Path(p) if p else pathlib.Path.home() / ".cache" / "rfnetwork_amd" / "launch.json"


_CFG_KEYS = ("variant", "block", "x_planes")
_build_id = None


def build_id():
    """The library's version and a digest of the built library file: any rebuild with changed
    kernels or a changed config format gives new keys, so no entry tuned for other kernels is
    applied."""
    global _build_id
    if _build_id is None:
        try:
            digest = hashlib.sha1(pathlib.Path(_lib.lib()._name).read_bytes()).hexdigest()[:12]
        except (OSError, AttributeError):
            digest = "nofile"
        _build_id = f"v{_lib.version()}-{digest}"
    return _build_id


def shape_key(grids, dtype, coef_kind):
    """grids: [(Nx, Ny, Nz, K)]; the key also holds the build id (kernels change with it)."""
    g = ";".join(f"{nx}x{ny}x{nz}x{k}" for nx, ny, nz, k in grids)
    return f"{build_id()}|{dtype}|coef{coef_kind}|{g}"


def _valid(cfg):
    return isinstance(cfg, dict) and all(isinstance(cfg.get(k), int) for k in _CFG_KEYS)


def _load():
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(key, cfg):
    try:
        path = cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        data = _load()
        data[key] = cfg
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass


def trial_steps_for(grids, requested=20, limit=None):
    """Steps per tuner trial: `requested`, more for small grids, whose steps take a few microseconds
    (about 3 ms of work per trial at 50 M cell-updates per ms, at most 200 steps), and fewer above
    50 M cells, whose steps take milliseconds (the trials' cost grows with the grid, their timing
    error does not; at least 6), at most `limit` (the installed series' length)."""
    cells = max(sum(nx * ny * nz * k for nx, ny, nz, k in grids), 1)
    base = int(requested) if cells <= 5e7 else max(6, int(requested * 5e7 / cells))
    n = max(base, min(200, int(1.5e8 // cells)))
    return min(n, int(limit)) if limit else n


def run_tuner(session, trial_steps=20):
    """Run the engine's tuner on a Session or Batch that has not stepped yet; the best shape is
    left applied.  Returns {variant, block, x_planes (0 = automatic), order (the separate kernels'
    workgroup numbering), ms_per_step}."""
    cfg = (ctypes.c_int32 * 3)()
    ms = ctypes.c_double()
    check(session._L.rfdtd_autotune(session._h, int(trial_steps), ctypes.cast(cfg, ctypes.c_void_p),
                                    ctypes.byref(ms)), "rfdtd_autotune")
    return dict(variant=int(cfg[0]), block=int(cfg[1]), x_planes=int(cfg[2]), order=session.column_order(),
                ms_per_step=float(ms.value))


def apply(session, cfg):
    session.set_variant(cfg["variant"])
    session.set_launch(cfg["block"], cfg["x_planes"] if cfg["x_planes"] > 0 else -1, 0)
    session.set_column_order(cfg.get("order", 0))  # entries without it were tuned with order 0


def _adjoint_from_cache(session, key, use_cache):
    """Apply the reverse-pass shape cached for key (timed by an earlier gradient), if any; later
    gradients of the session record theirs under key (remember_adjoint)."""
    session._tune_key = key if use_cache else None
    if not use_cache:
        return
    akey = key + "|adjoint"
    acfg = _memory.get(akey) or _load().get(akey)
    if acfg:
        _memory[akey] = acfg
        session.set_adjoint_launch(acfg["variant"], acfg["block"], acfg["x_chunks"])


def remember_adjoint(session):
    """Cache the reverse-pass shape that a gradient of an autotuned session timed."""
    key = getattr(session, "_tune_key", None)
    if key is None:
        return
    cfg = session.adjoint_launch()
    akey = key + "|adjoint"
    if cfg["block"] > 0 and _memory.get(akey) != cfg:
        _memory[akey] = cfg
        _store(akey, cfg)


def tune(session, grids, dtype, coef_kind, trial_steps=20, use_cache=True):
    """Tuned launch shape of the session's grids: from the cache when known, else measured (and
    cached).  The shape is applied to the session either way, with the reverse-pass shape when one
    is cached.  Returns the config with cached=True/False."""
    key = shape_key(grids, dtype, coef_kind)
    if getattr(session, "flags", 0) & _lib.TUNE_SEPARATE:  # a restricted candidate set tunes apart
        key += "|separate"
    if getattr(session, "_history", False):  # so does a session recording a design-box history
        key += "|history"
    trial_steps = trial_steps_for(grids, trial_steps, getattr(session, "Nt", None))
    _adjoint_from_cache(session, key, use_cache)
    cfg = _memory.get(key)
    if cfg is None and use_cache:
        cfg = _load().get(key)
    if not _valid(cfg):
        cfg = None  # an entry of another format counts as a miss
    if cfg is not None and use_cache:
        apply(session, cfg)
        _memory[key] = cfg
        return dict(cfg, cached=True)
    cfg = run_tuner(session, trial_steps)
    _memory[key] = cfg
    if use_cache:
        _store(key, cfg)
    return dict(cfg, cached=False)
