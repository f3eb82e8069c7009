"""Seeded synthetic inputs of the five benchmark workloads, shared by
``bench.py`` and the full-size GPU tests so that both run the same atoms.

  trajectory   100,000 atoms, 100 A box, H/C/O/N/S 50/30/15/4/1 %, 100 frames
               (seed structure + N(0, 0.5^2 A^2) per frame), 128^3 cell,
               sigma 2, q <= 0.6 with 0.005 bins (reset to the grid spacing)
  uneven       32 systems, 10k-300k atoms at 0.1 atoms/A^3, cubic grids of
               the smallest 2,3,5,7-smooth size >= L / 0.75 A in [64, 192],
               sigma 2, 8 frames each
  restraint    50,000 atoms, 80 A, 128^3, sigma 2, order 6, chi^2 target from
               the structure displaced by N(0, 1 A^2), 100 bins 0.005-0.5
  large        1,000,000 atoms, 215 A, 9 elements (H 48 %, C 30, O 15, N 4,
               S 1, P/Na/Cl/K 0.5 each), 256^3 cell, sigma 2, 16 frames
  single       5,000 atoms, 37 A, 64^3, sigma 2 (the 5k-atom frame)

The reference ships no benchmark inputs; its documented measurements use
MDAnalysis trajectories (docs/algorithm/performance.md:176-230).
"""
from __future__ import annotations

import numpy as np

MIX5 = (("H", 0.50), ("C", 0.30), ("O", 0.15), ("N", 0.04), ("S", 0.01))
MIX9 = (("H", 0.48), ("C", 0.30), ("O", 0.15), ("N", 0.04), ("S", 0.01), ("P", 0.005),
        ("Na", 0.005), ("Cl", 0.005), ("K", 0.005))
DISP = 0.5

TRAJ = dict(n_atoms=100_000, box=100.0, grid=128, sigma=2.0, dq=0.005, qcut=0.6, order=4,
            frames=100, seed=2)


def smooth_at_least(x: float) -> int:
    """Smallest 2,3,5,7-smooth integer >= x."""
    v = max(1, int(np.ceil(x)))
    while True:
        m = v
        for p_ in (2, 3, 5, 7):
            while m % p_ == 0:
                m //= p_
        if m == 1:
            return v
        v += 1


def mix_symbols(rng, n_atoms, mix):
    """Exact element counts of the mix (rounding remainder on the first
    element), randomly permuted; returns (sorted symbols, element id per atom)."""
    counts = {s_: int(round(n_atoms * f)) for s_, f in mix}
    counts[mix[0][0]] += n_atoms - sum(counts.values())
    syms = []
    for s_, c in counts.items():
        syms += [s_] * c
    syms = np.asarray(syms)[rng.permutation(n_atoms)]
    symbols = sorted(set(syms.tolist()))
    lut = {s_: i for i, s_ in enumerate(symbols)}
    return symbols, np.array([lut[s_] for s_ in syms], dtype=np.int32)


def trajectory(n_atoms=TRAJ["n_atoms"], frames=TRAJ["frames"]):
    """Positions [frames, n_atoms, 3] float32, sorted symbols, element ids and
    the per-atom symbol list of the 100k-atom trajectory (base seed 2, frame
    stream default_rng([2, frame]))."""
    seed, box = TRAJ["seed"], TRAJ["box"]
    rng = np.random.default_rng(seed)
    base = rng.uniform(0.0, box, size=(n_atoms, 3)).astype(np.float32)
    counts = {s: int(round(n_atoms * f)) for s, f in MIX5}
    counts["H"] += n_atoms - sum(counts.values())
    syms = []
    for s, c in counts.items():
        syms += [s] * c
    syms = list(np.asarray(syms)[rng.permutation(n_atoms)])
    xyz = np.empty((frames, n_atoms, 3), dtype=np.float32)
    for f in range(frames):
        r = np.random.default_rng([seed, f])
        xyz[f] = base + r.normal(0.0, DISP, size=base.shape).astype(np.float32)
    symbols = sorted(set(syms))
    lut = {s: i for i, s in enumerate(symbols)}
    ids = np.array([lut[s] for s in syms], dtype=np.int32)
    return xyz, symbols, ids, syms


def uneven_systems(count=32, frames=8, seed=3):
    """The 32 uneven systems: dicts with n_atoms, L (float32 value), n,
    symbols, ids and xyz [frames, n_atoms, 3]."""
    out = []
    rng = np.random.default_rng(seed)
    for s_ in range(count):
        n_atoms = int(round(np.exp(rng.uniform(np.log(1e4), np.log(3e5)))))
        L = (n_atoms / 0.1) ** (1.0 / 3.0)
        n = min(192, max(64, smooth_at_least(L / 0.75)))
        r = np.random.default_rng([seed, s_])
        base = r.uniform(0.0, L, size=(n_atoms, 3)).astype(np.float32)
        symbols, ids = mix_symbols(r, n_atoms, MIX5)
        xyz = np.stack([base + np.random.default_rng([seed, s_, f]).normal(0.0, DISP, size=base.shape)
                        .astype(np.float32) for f in range(frames)])
        out.append(dict(n_atoms=n_atoms, L=float(np.float32(L)), n=n, symbols=symbols, ids=ids,
                        xyz=xyz))
    return out


def restraint():
    """50,000 atoms in an 80 A box, 128^3 cell, sigma 2, order 6; target
    structure displaced by N(0, 1 A^2); chi^2 over the requested 100 bins from
    0.005 to 0.5 A^-1 (exact bins) with 1 % errors."""
    rng = np.random.default_rng(4)
    n_atoms, L = 50_000, 80.0
    base = rng.uniform(0.0, L, size=(n_atoms, 3)).astype(np.float32)
    symbols, ids = mix_symbols(rng, n_atoms, MIX5)
    target = (base + np.random.default_rng([4, 1]).normal(0.0, 1.0, size=base.shape)
              ).astype(np.float32)
    return dict(base=base, target=target, symbols=symbols, ids=ids, L=L, n=128, order=6,
                dq=0.005, qcut=0.5)


LARGE = dict(n_atoms=1_000_000, box=215.0, grid=256, frames=16, qcut=0.5, seed=5)


def large_base():
    """Seed structure, sorted symbols and element ids of the 1M-atom system."""
    rng = np.random.default_rng(LARGE["seed"])
    base = rng.uniform(0.0, LARGE["box"], size=(LARGE["n_atoms"], 3)).astype(np.float32)
    symbols, ids = mix_symbols(rng, LARGE["n_atoms"], MIX9)
    return base, symbols, ids


def large_frame(base, f):
    """Frame f of the 1M-atom system (seed structure + N(0, 0.5^2 A^2))."""
    return base + np.random.default_rng([LARGE["seed"], f]).normal(
        0.0, DISP, size=base.shape).astype(np.float32)
