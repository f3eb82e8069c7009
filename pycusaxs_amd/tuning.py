"""Autotuner for the engine's launch shapes.

For one system (cell grid, FFT grid, order, element count, atom density)
it times every candidate spreading tile, z-transform variant (pruned radix-16
or full Stockham) and Stockham radix factorization (8/4/2 then 3, 5, 7, or
the widest register radices first) on synthetic atoms of the same density
with the engine's hipEvent hook, and keeps the fastest.  Results are cached per device name in a JSON file
(``PYCUSAXS_AMD_TUNE_CACHE``, default ~/.cache/pycusaxs_amd/tune.json) so
a shape is tuned once.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

from . import _lib

TILES = ((4, 8), (8, 4), (4, 4), (2, 8))
VARIANTS = (0, _lib.FLAG_NO_PRUNED_Z, _lib.FLAG_WIDE_RADIX)
TIMED = ("prep", "scan", "scatter", "spread_z", "yfft", "xfft_epi", "reduce")


def _cache_path() -> Path:
    p = os.environ.get("PYCUSAXS_AMD_TUNE_CACHE")
    if p:
        return Path(p)
    return Path.home() / ".cache" / "pycusaxs_amd" / "tune.json"


def shape_key(spec, n_atoms: int, box_volume: float, forward_only: bool = False) -> str:
    density = n_atoms / max(box_volume, 1e-9)
    dbucket = int(round(math.log2(max(density, 1e-6)) * 2))        # half-octaves
    return (f"n{'x'.join(map(str, spec.n))}-nn{'x'.join(map(str, spec.nn))}-p{spec.order}"
            f"-e{len(spec.symbols)}-q{spec.qcut_req:g}-d{dbucket}" + ("-fwd" if forward_only else ""))


class Autotuner:
    def __init__(self, cache_path: Optional[Path] = None, frames: int = 4, repeats: int = 2):
        self.path = Path(cache_path) if cache_path else _cache_path()
        self.frames = frames
        self.repeats = repeats
        self._cache: Dict[str, Dict] = {}
        if self.path.exists():
            try:
                self._cache = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self._cache = {}

    def _device(self) -> str:
        import torch
        return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"

    def lookup(self, spec, n_atoms: int, box,
               forward_only: bool = False) -> Optional[Tuple[Tuple[int, int], int]]:
        vol = float(np.prod(np.diag(np.asarray(box, dtype=np.float64).reshape(3, 3))))
        ent = self._cache.get(self._device(), {}).get(shape_key(spec, n_atoms, vol, forward_only))
        if ent is None:
            return None
        return tuple(ent["tile"]), int(ent["flags"])

    def tune(self, spec, n_atoms: int, box, candidates: Optional[Iterable] = None,
             seed: int = 0, forward_only: bool = False) -> Tuple[Tuple[int, int], int]:
        """Times the candidates (tile, flags) and returns the fastest.  With
        forward_only the plans are made with FLAG_FORWARD_ONLY (tiles fitted
        to the forward kernels alone; such plans refuse backward calls)."""
        import torch
        from .engine import SaxsEngine
        hit = self.lookup(spec, n_atoms, box, forward_only)
        if hit is not None:
            return hit
        b = np.diag(np.asarray(box, dtype=np.float64).reshape(3, 3))
        rng = np.random.default_rng(seed)
        xyz = torch.from_numpy((rng.uniform(0, 1, size=(self.frames, n_atoms, 3)) * b)
                               .astype(np.float32)).cuda()
        elem = torch.from_numpy(rng.integers(0, len(spec.symbols), size=n_atoms)
                                .astype(np.int32)).cuda()
        cands = list(candidates) if candidates is not None else \
            [(t, f) for t in TILES for f in VARIANTS]
        if forward_only:
            cands = [(t, f | _lib.FLAG_FORWARD_ONLY) for t, f in cands]
        results = []
        for tile, flags in cands:
            try:
                eng = SaxsEngine(spec, box, max_atoms=n_atoms, max_frames=self.frames,
                                 tile=tile, flags=flags)
            except (ValueError, RuntimeError):
                continue
            eng.forward(xyz, elem)
            torch.cuda.synchronize()
            eng.set_timing(True)
            for _ in range(self.repeats):
                eng.forward(xyz, elem)
            torch.cuda.synchronize()
            t = eng.kernel_times()
            eng.close()
            total = sum(ms for k, (ms, _) in t.items() if k in TIMED)
            results.append((total, tile, flags))
        if not results:
            raise RuntimeError("autotuner: no candidate launch shape is valid for this system")
        results.sort()
        best_ms, tile, flags = results[0]
        vol = float(np.prod(b))
        dev = self._cache.setdefault(self._device(), {})
        dev[shape_key(spec, n_atoms, vol, forward_only)] = {
            "tile": list(tile), "flags": int(flags), "ms_per_frame": best_ms / (self.repeats * self.frames),
            "candidates": [[list(t), int(f), ms / (self.repeats * self.frames)] for ms, t, f in results]}
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            self.path.write_text(json.dumps(self._cache, indent=1))
        except OSError:
            pass
        return tuple(tile), int(flags)
