"""Writers of GROMACS TRR, XTC and GRO files for the reader's tests.

No trajectory file exists in the reference or in this image, so the tests
write their own from the published formats (the reader's header in
pycusaxs_amd/csrc/xdr_frames.cpp states them).  The XTC encoder is the
packing side written independently of the decoder: it chooses its own
groups (a full atom, a run of small-difference atoms whose first member
precedes the full atom, kept run lengths sent as a 0 flag, level changes up
and down, the separate-field path for ranges above 2^24), and decoding must
give back its integers exactly.
"""
import struct

import numpy as np

# XTC small-difference levels (format constants; index = bits per triplet)
MAGIC = [0] * 9 + [
    8, 10, 12, 16, 20, 25, 32, 40, 50, 64, 80, 101, 128, 161, 203, 256, 322, 406, 512, 645,
    812, 1024, 1290, 1625, 2048, 2580, 3250, 4096, 5060, 6501, 8192, 10321, 13003, 16384,
    20642, 26007, 32768, 41285, 52015, 65536, 82570, 104031, 131072, 165140, 208063, 262144,
    330280, 416127, 524287, 660561, 832255, 1048576, 1321122, 1664510, 2097152, 2642245,
    3329021, 4194304, 5284491, 6658042, 8388607, 10568983, 13316085, 16777216]
FIRST = 9
LAST = len(MAGIC) - 1


def _i(v):
    return struct.pack(">i", int(v))


def _f(v):
    return struct.pack(">f", float(v))


def _d(v):
    return struct.pack(">d", float(v))


def trr_frame(x_nm, box_nm, step, t, double=False, v=None, f=None, lam=0.0):
    n = len(next(b for b in (x_nm, v, f) if b is not None))
    w = 8 if double else 4
    real = _d if double else _f
    out = _i(1993) + _i(13) + _i(12) + b"GMX_trn_file"
    sizes = [0, 0, 9 * w if box_nm is not None else 0, 0, 0, 0, 0,
             3 * n * w if x_nm is not None else 0, 3 * n * w if v is not None else 0,
             3 * n * w if f is not None else 0]
    for s in sizes:
        out += _i(s)
    out += _i(n) + _i(step) + _i(0) + real(t) + real(lam)
    for blk in (box_nm, x_nm, v, f):
        if blk is not None:
            out += b"".join(real(val) for val in np.asarray(blk, dtype=np.float64).reshape(-1))
    return out


class _Bits:
    def __init__(self):
        self.bits = []

    def put(self, value, n):
        assert 0 <= value < (1 << n) or n == 0, (value, n)
        for k in range(n - 1, -1, -1):
            self.bits.append((value >> k) & 1)

    def put_triplet(self, vals, sizes, nbits):
        a0, a1, a2 = (int(v) for v in vals)
        assert all(0 <= a < s for a, s in zip(vals, sizes)), (vals, sizes)
        num = (a0 * sizes[1] + a1) * sizes[2] + a2
        while nbits > 8:
            self.put(num & 0xFF, 8)
            num >>= 8
            nbits -= 8
        self.put(num, nbits)

    def tobytes(self):
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(int("".join(map(str, b[k:k + 8])), 2) for k in range(0, len(b), 8))


def _fits(d, idx):
    half = MAGIC[idx] // 2
    return all(0 <= int(c) + half < MAGIC[idx] for c in d)


def xtc_coords(ints, precision, rng, start_idx=12, max_small=8, stats=None):
    """Packed coordinate block of one frame (natoms > 9) from integer
    coordinates ints [n, 3] (nm times precision).  `stats` (a dict) counts
    the packing choices made: runs, kept run lengths, level changes."""
    stats = stats if stats is not None else {}
    for key in ("runs", "kept", "up", "down", "longest"):
        stats.setdefault(key, 0)
    ints = np.asarray(ints, dtype=np.int64)
    n = len(ints)
    mn = ints.min(axis=0)
    mx = ints.max(axis=0)
    sizes = [int(v) for v in (mx - mn + 1)]
    large = any(s > 0xFFFFFF for s in sizes)
    if large:
        fbits = [int(s).bit_length() for s in sizes]
    else:
        bitsize = (sizes[0] * sizes[1] * sizes[2]).bit_length()
    bw = _Bits()
    idx = start_idx
    prev_run = 0
    i = 0
    while i < n:
        # group: atom i as the run's first member before full atom i + 1,
        # then atoms i + 2, ... each relative to the previous run member
        small = []
        if i + 1 < n and _fits(ints[i] - ints[i + 1], idx):
            small = [i]
            j = i + 2
            ref = i
            while j < n and len(small) < max_small and _fits(ints[j] - ints[ref], idx):
                small.append(j)
                ref = j
                j += 1
        full = i + 1 if small else i
        choices = [0] + ([-1] if idx > FIRST else []) + ([1] if idx < LAST else [])
        change = int(rng.choice(choices))
        if 3 * len(small) + change + 1 > 31:
            change = 0 if 3 * len(small) + 1 <= 31 else -1
        v = ints[full] - mn
        if large:
            for c in range(3):
                bw.put(int(v[c]), fbits[c])
        else:
            bw.put_triplet(v, sizes, bitsize)
        run = 3 * len(small)
        if change == 0 and run == prev_run and rng.random() < 0.5:
            bw.put(0, 1)                                 # run length kept
            stats["kept"] += 1
        else:
            bw.put(1, 1)
            bw.put(run + change + 1, 5)
        prev_run = run
        stats["runs"] += 1 if small else 0
        stats["longest"] = max(stats["longest"], len(small))
        stats["up"] += change > 0
        stats["down"] += change < 0
        half = MAGIC[idx] // 2
        ref = full
        for k, a in enumerate(small):
            d = ints[a] - ints[ref]
            bw.put_triplet([int(c) + half for c in d], [MAGIC[idx]] * 3, idx)
            ref = a
        i = max([full] + small) + 1
        idx += change
    data = bw.tobytes()
    out = _f(precision) + b"".join(_i(v) for v in mn) + b"".join(_i(v) for v in mx)
    out += _i(start_idx) + _i(len(data)) + data + b"\0" * (-len(data) % 4)
    return out


def xtc_frame(ints_or_nm, box_nm, step, t, precision=1000.0, rng=None, start_idx=12,
              max_small=8, stats=None):
    """One XTC frame.  For n > 9 atoms ints_or_nm are the integer
    coordinates; for n <= 9 the floats (nm) stored uncompressed."""
    n = len(ints_or_nm)
    out = _i(1995) + _i(n) + _i(step) + _f(t)
    out += b"".join(_f(v) for v in np.asarray(box_nm, dtype=np.float64).reshape(-1))
    out += _i(n)
    if n <= 9:
        out += b"".join(_f(v) for v in np.asarray(ints_or_nm, dtype=np.float64).reshape(-1))
    else:
        rng = rng if rng is not None else np.random.default_rng(0)
        out += xtc_coords(ints_or_nm, precision, rng, start_idx, max_small, stats)
    return out


def quantize(x_nm, precision):
    """GROMACS' rounding of coordinates to integers (half away from zero)."""
    v = np.asarray(x_nm, dtype=np.float64) * float(np.float32(precision))
    return np.where(v >= 0, np.floor(v + 0.5), np.ceil(v - 0.5)).astype(np.int64)


def expected_angstrom(ints, precision):
    """What the decoder returns for integer coordinates: float32 int times
    the float32 reciprocal of the precision, times 10."""
    inv = np.float32(1.0 / float(np.float32(precision)))
    return (np.asarray(ints).astype(np.float32) * inv) * np.float32(10.0)


def water_like(n_mol, box_nm, rng, spread=0.06):
    """Clustered triplets (O, H, H about `spread` nm apart) spread over the
    box: the shape XTC's runs are made for."""
    o = rng.uniform(0.0, box_nm, size=(n_mol, 3))
    h1 = o + rng.normal(0, spread, size=(n_mol, 3))
    h2 = o + rng.normal(0, spread, size=(n_mol, 3))
    return np.stack([o, h1, h2], axis=1).reshape(-1, 3)


def gro_text(names, resnames, x_nm, box_nm, title="test system"):
    lines = [title, "%5d" % len(names)]
    for i, (a, r, p) in enumerate(zip(names, resnames, x_nm)):
        lines.append("%5d%-5s%5s%5d%8.3f%8.3f%8.3f" % (i // 3 + 1, r, a, (i + 1) % 100000,
                                                       p[0], p[1], p[2]))
    b = np.asarray(box_nm, dtype=np.float64)
    if np.count_nonzero(b - np.diag(np.diag(b))) == 0:
        lines.append("%10.5f%10.5f%10.5f" % (b[0, 0], b[1, 1], b[2, 2]))
    else:
        lines.append("%10.5f" * 9 % (b[0, 0], b[1, 1], b[2, 2], b[0, 1], b[0, 2], b[1, 0],
                                      b[1, 2], b[2, 0], b[2, 1]))
    return "\n".join(lines) + "\n"
