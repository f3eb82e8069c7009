"""Cromer-Mann atomic form factors: f(q) = sum_l a_l exp(-b_l (q/4pi)^2) + c.

The coefficient table (International Tables for Crystallography, 1992
edition, as tabulated in Scattering.cpp:9-189; species "Du" is a unit
dummy) is stored as data in data/cromer_mann_it1992.json:
symbol -> [a1, a2, a3, a4, b1, b2, b3, b4, c].
"""
from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path
from typing import Dict, Tuple

_DATA = Path(__file__).resolve().parent / "data" / "cromer_mann_it1992.json"


@lru_cache(maxsize=1)
def table() -> Dict[str, Tuple[float, ...]]:
    with open(_DATA) as fh:
        raw = json.load(fh)
    return {k: tuple(float(x) for x in v) for k, v in raw.items()}


def coefficients(symbol: str) -> Tuple[float, ...]:
    """Nine coefficients of a species; unknown symbols raise ValueError
    (the reference reads an empty entry there, Scattering.h:15)."""
    t = table()
    if symbol not in t:
        raise ValueError(f"no atomic form factor for element {symbol!r}")
    return t[symbol]
