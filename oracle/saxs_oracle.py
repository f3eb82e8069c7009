"""FP64 CPU oracle of the pyCuSaxs per-frame SAXS scattering path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``, the
smoke check of the repository's build entry module and the CPU-baseline leg
of ``bench.py`` may import it.  The product (``pycusaxs_amd``) never imports
or calls it.

It restates the reference's algorithm literally, in float64 where the
reference computes in float32, keeping the reference's float32 steps only
where they decide integers (grid cell of an atom, FFT size, bin parameters):

  setup   find_closest_product        saxsKernel.cu:377-469
          hist_params (dq/qcut reset) saxsKernel.cu:489-526, :247
          border_planes               saxsKernel.h:107-110, saxsKernel.cu:68-70
          bsp_moduli                  BSpmod.cpp:45-129, BSpmod.h:22
          inverse_box(_matrix)        Cell.cpp:111-126 (applied to the transposed box)
  frame   spread (S1)                 saxsDeviceKernels.cu:338-398, Splines.cu:11-74
          pad (S2)                    saxsDeviceKernels.cu:446-465
          pad, "given" mode           Options.h:7-11, AtomCounter.cpp:56-103
          embed (S3)                  saxsDeviceKernels.cu:413-429
          R2C + form factor (S4)      saxsKernel.cu:171, saxsDeviceKernels.cu:177-216,
                                      opsfact.cu:29-35
          modulus (S5)                saxsDeviceKernels.cu:142-162
          histogram (S6)              saxsDeviceKernels.cu:67-125
  NVT     spectrum sum, one binning   saxsKernel.cu:17-39,184-189, RunSaxs.cu:124-130,208-211
  output  rows (b*dq, H_b/C_b)        saxsKernel.cu:210-225, RunSaxs.cu:224-227

Where the reference's float32 rounding at exact bin edges is undefined
(integer |m| shells falling exactly on b*dq), bins follow the exact rule
``bin(n) = max{b : b^2 dq^2 <= n s^2}`` with membership ``n s^2 <= qcut^2``
evaluated on the float32 values as exact rationals (``fractions``), where
``n = |m|^2`` and ``s`` is the reciprocal grid spacing.  The engine's host
code implements the same rule independently (integer arithmetic); the tests
pin both against brute-force enumeration and the known shell counts.

Pinning (see tests/test_oracle_*.py): B-spline weights and moduli against
closed forms and the compiled reference helpers (oracle/_ref); the FFT
against a brute-force DFT; spreading against single-atom stencils and
partition of unity; end-to-end I(m) against the direct structure factor for
atoms on grid nodes; bin counts against the known answers 34, 82, 178, 258,
402 (64^3 cell, 128^3 FFT, 37 A box); the border-slab pad against its closed
form (the z-plane B-spline weights of atoms straddling the slab edges, and
a full lattice whose density is exactly 1); the dq/qcut reset, row count
and q column against the five profiles stored in the reference's own
database (pycusaxs/data/reference_solvents.db, two of them with a lowered
cutoff).  The adjoint is pinned by central finite differences of the
forward oracle.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

F32 = np.float32

# FFT used by the oracle: numpy's pocketfft by default; the CPU baseline in
# bench.py may switch it to scipy.fft with several worker threads.
_rfftn = np.fft.rfftn
_ifftn = np.fft.ifftn


def set_fft_workers(workers: int) -> int:
    """Use scipy.fft with this many threads (1 = numpy); returns the number
    of threads in effect."""
    global _rfftn, _ifftn
    if workers <= 1:
        _rfftn, _ifftn = np.fft.rfftn, np.fft.ifftn
        return 1
    import scipy.fft as sfft
    _rfftn = lambda a: sfft.rfftn(a, workers=workers)  # noqa: E731
    _ifftn = lambda a: sfft.ifftn(a, workers=workers)  # noqa: E731
    return workers

# Cromer-Mann coefficients {a1..a4, b1..b4, c} for the species the tests use,
# transcribed independently of the engine's data file from
# Scattering.cpp:10 (Du), :11 (H), :20 (C), :22 (N), :23 (O), :30 (Na),
# :41 (P), :42 (S), :43 (Cl), :46 (K).
FORM_FACTORS: Dict[str, Tuple[float, ...]] = {
    "Du": (1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
    "H": (0.493002, 0.322912, 0.140191, 0.040810, 10.5109, 26.1257, 3.14236, 57.7997, 0.003038),
    "C": (2.31000, 1.02000, 1.58860, 0.865000, 20.8439, 10.2075, 0.568700, 51.6512, 0.215600),
    "N": (12.2126, 3.13220, 2.01250, 1.16630, 0.005700, 9.89330, 28.9975, 0.582600, -11.529),
    "O": (3.04850, 2.28680, 1.54630, 0.867000, 13.2771, 5.70110, 0.323900, 32.9089, 0.250800),
    "Na": (4.76260, 3.17360, 1.26740, 1.11280, 3.28500, 8.84220, 0.313600, 129.424, 0.676000),
    "P": (6.43450, 4.17910, 1.78000, 1.49080, 1.90670, 27.1570, 0.526000, 68.1645, 1.11490),
    "S": (6.90530, 5.20340, 1.43790, 1.58630, 1.46790, 22.2151, 0.253600, 56.1720, 0.866900),
    "Cl": (11.4604, 7.19640, 6.25560, 1.64550, 0.010400, 1.16620, 18.5194, 47.7784, -9.5574),
    "K": (8.21860, 7.43980, 1.05190, 0.865900, 12.7949, 0.774800, 213.187, 41.6841, 1.42280),
}

SHELL = F32(0.07)          # Parameters.h:8
GAMMA_TERMS = 50           # BSpmod.h:22 (KCUT)
MODULI_TINY = 1.0e-7       # BSpmod.cpp:104


# ----------------------------------------------------------------- setup ---

def generate_multiples(limit: int) -> List[int]:
    """All 2^a 3^b 5^c 7^d <= limit, sorted (saxsKernel.cu:377-436)."""
    out = set()
    a = 1
    while a <= limit:
        b = a
        while b <= limit:
            c = b
            while c <= limit:
                d = c
                while d <= limit:
                    out.add(d)
                    d *= 7
                c *= 5
            b *= 3
        a *= 2
    return sorted(out)


def find_closest_product(n: int, sigma: float) -> int:
    """7-smooth size nearest round(n*sigma); ties go to the smaller one
    (saxsKernel.cu:449-469).  n*sigma is a float32 product; std::round
    rounds halves away from zero."""
    x = float(F32(n) * F32(sigma))
    target = int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))
    closest, best = target, None
    for m in generate_multiples(2 * target):
        d = abs(m - target)
        if best is None or d < best:
            best, closest = d, m
    return closest


def border_planes(nz: int) -> int:
    """d = int(0.07f * nz / 2) in float32 (saxsKernel.h:107-110)."""
    return int(F32(F32(SHELL * F32(nz)) / F32(2.0)))


def inverse_box(box) -> np.ndarray:
    """Diagonal of oc for an orthorhombic box: float32(1.0 / box_aa)
    (Cell.cpp:118-125; the off-diagonal terms vanish for a diagonal box)."""
    box = np.asarray(box, dtype=np.float32).reshape(3, 3)
    return np.array([F32(1.0 / float(box[a, a])) for a in range(3)], dtype=np.float32)


def inverse_box_matrix(box) -> np.ndarray:
    """M with fractional coordinates u = M x for a box whose rows are the cell
    vectors a = (a_x, 0, 0), b = (b_x, b_y, 0), c = (c_x, c_y, c_z) (the
    MDAnalysis/GROMACS lower-triangular form), so x = B^T u and M = (B^T)^-1.
    B^T is upper-triangular, whose inverse is the formula of Cell.cpp:118-125
    (float32 products, 1.0/x in double): applied to B^T it gives the true
    inverse, where the reference applies it to B itself and loses the
    off-diagonal terms (Cell.cpp:111-126)."""
    B = np.asarray(box, dtype=np.float32).reshape(3, 3)
    if B[0, 1] != 0 or B[0, 2] != 0 or B[1, 2] != 0:
        raise ValueError("box must be lower-triangular (rows a, b, c with a || x, b in xy)")
    co = B.T
    M = np.zeros((3, 3), dtype=np.float32)
    M[0, 0] = F32(1.0 / float(co[0, 0]))
    M[0, 1] = F32(-co[0, 1] / F32(co[0, 0] * co[1, 1]))
    M[0, 2] = F32(F32(F32(co[0, 1] * co[1, 2]) - F32(co[0, 2] * co[1, 1]))
                  / F32(F32(co[0, 0] * co[1, 1]) * co[2, 2]))
    M[1, 1] = F32(1.0 / float(co[1, 1]))
    M[1, 2] = F32(-co[1, 2] / F32(co[1, 1] * co[2, 2]))
    M[2, 2] = F32(1.0 / float(co[2, 2]))
    return M


def reciprocal_basis(M, n, nn) -> np.ndarray:
    """Rows b_a (float32) with q = sum_a m_a b_a for FFT index m: the
    transform of a function of u = M x at frequency m_a n_a / nn_a per cell is
    exp(-i q.x) with q_j = 2 pi sum_a M_aj m_a n_a / nn_a.  Each entry is
    evaluated like the spacing of hist_params, so a diagonal M gives
    diag(s_a) bit for bit."""
    tp = F32(F32(2.0) * F32(math.pi))
    b = np.zeros((3, 3), dtype=np.float32)
    for a in range(3):
        ratio = F32(F32(nn[a]) / F32(n[a]))
        for j in range(3):
            b[a, j] = F32(F32(tp * F32(M[a, j])) / ratio)
    return b


@dataclass
class HistParams:
    dq: np.float32          # effective bin width (float32, saxsKernel.cu:506)
    qcut: np.float32        # effective cutoff (float32, saxsKernel.cu:507)
    num_bins: int           # int(qcut/dq) + 1 (saxsKernel.cu:247)
    spacing: Tuple[np.float32, np.float32, np.float32]   # reciprocal spacing s_a


def hist_params(n, nn, oc_diag, sigma, dq_req, qcut_req, dq_policy="reference") -> HistParams:
    """dq/qcut reset of saxsKernel.cu:489-526, all in float32.

    arg_a = 2pi_f * oc_aa / sigma (the reference's grid spacing, which uses
    sigma even when nn/n differs: Q2).  dq = max(arg_x, arg_y, arg_z, dq_req),
    qcut = min(arg_a*(nf_a-1), qcut_req).  With dq_policy="exact" the
    requested dq is kept (qcut is still clipped to the grid).
    The point spacing s_a uses the actual ratio nn_a/n_a in the same float32
    expression, so s_a == arg_a bit-for-bit whenever nn_a/n_a == sigma.
    """
    two_pi = F32(F32(2.0) * F32(math.pi))
    sig = F32(sigma)
    arg = [F32(F32(two_pi * F32(oc_diag[a])) / sig) for a in range(3)]
    nf = [x // 2 if x % 2 == 0 else x // 2 + 1 for x in nn]
    cut = [F32(arg[a] * F32(nf[a] - 1)) for a in range(3)]
    dq = max(arg + [F32(dq_req)]) if dq_policy == "reference" else F32(dq_req)
    qcut = min(cut + [F32(qcut_req)])
    num_bins = int(F32(qcut / dq)) + 1
    s = tuple(F32(F32(two_pi * F32(oc_diag[a])) / F32(F32(nn[a]) / F32(n[a]))) for a in range(3))
    return HistParams(F32(dq), F32(qcut), num_bins, s)


def bin_of_n(nsq: int, s: np.float32, dq: np.float32, qcut: np.float32, num_bins: int) -> int:
    """Exact bin of the shell n = |m|^2: -1 outside q <= qcut or b >= num_bins."""
    S = Fraction(float(s))
    return bin_of_q2(nsq * S * S, dq, qcut, num_bins)


def bin_of_m(m, spacing, dq, qcut, num_bins: int) -> int:
    """Exact bin of the frequency m on an orthorhombic reciprocal grid:
    q^2 = sum_a (m_a s_a)^2 as an exact rational of the float32 spacings
    (the reference evaluates 2 pi oc_aa m_a per axis, saxsDeviceKernels.cu:84-91)."""
    q2 = sum(int(m[a]) ** 2 * Fraction(float(spacing[a])) ** 2 for a in range(3))
    return bin_of_q2(q2, dq, qcut, num_bins)


def bin_of_vec(m, basis, dq, qcut, num_bins: int) -> int:
    """Exact bin of frequency m on a triclinic reciprocal grid: q = sum_a m_a b_a
    with the float32 rows b_a as exact rationals."""
    q2 = Fraction(0)
    for j in range(3):
        qj = sum(int(m[a]) * Fraction(float(basis[a, j])) for a in range(3))
        q2 += qj * qj
    return bin_of_q2(q2, dq, qcut, num_bins)


def bin_of_q2(q2: Fraction, dq, qcut, num_bins: int) -> int:
    """b = max{b : b^2 dq^2 <= q^2}, -1 if q^2 > qcut^2 or b >= num_bins."""
    D, Q = Fraction(float(dq)), Fraction(float(qcut))
    if q2 > Q * Q:
        return -1
    b = int(math.isqrt(int(q2 / (D * D))))
    while (b + 1) ** 2 * D * D <= q2:
        b += 1
    while b > 0 and b * b * D * D > q2:
        b -= 1
    return b if b < num_bins else -1


def bspline_theta(w, p: int) -> np.ndarray:
    """B-spline weights by the reference recursion (Splines.cu:47-110,
    BSpline.cpp:281-316), in float64.  theta[k] = M_p(w + p - 1 - k)."""
    w = np.asarray(w, dtype=np.float64)
    th = [np.zeros_like(w) for _ in range(p)]
    th[0] = 1.0 - w
    th[1] = w.copy()
    for k in range(3, p + 1):
        div = 1.0 / (k - 1)
        th[k - 1] = div * w * th[k - 2]
        for j in range(1, k - 1):
            th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1])
        th[0] = div * (1.0 - w) * th[0]
    return np.stack(th, axis=-1)


def bspline_dtheta(w, p: int) -> np.ndarray:
    """d theta / d w: Diff() of the order p-1 weights (Splines.cu:84-92)."""
    t = bspline_theta(w, p - 1)
    d = np.zeros(t.shape[:-1] + (p,), dtype=np.float64)
    d[..., 0] = -t[..., 0]
    for o in range(1, p - 1):
        d[..., o] = t[..., o - 1] - t[..., o]
    d[..., p - 1] = t[..., p - 2]
    return d


def bsp_moduli(N: int, p: int) -> np.ndarray:
    """mod[k] = lambda(k)^2 / |b(k)|^2 (BSpmod.cpp:128-211), float64."""
    th0 = bspline_theta(0.0, p)
    A = np.zeros(p + 1)
    A[1:p + 1] = th0
    k = np.arange(N)
    t = np.arange(p + 1)
    arg = 2.0 * np.pi * np.outer(k, t) / N
    bsp = (A * np.cos(arg)).sum(1) ** 2 + (A * np.sin(arg)).sum(1) ** 2
    for o in range(1, N - 1):                      # BSpmod.cpp:167-168, sequential
        if not bsp[o] > MODULI_TINY:
            bsp[o] = 0.5 * (bsp[o - 1] + bsp[o + 1])
    nf = N // 2
    for kk in range(N):
        m = kk if kk < nf else kk - N
        if m != 0:
            x = math.pi * m / N
            g1 = 1.0 + sum((x / (x + math.pi * j)) ** p + (x / (x - math.pi * j)) ** p
                           for j in range(1, GAMMA_TERMS + 1))
            g2 = 1.0 + sum((x / (x + math.pi * j)) ** (2 * p) + (x / (x - math.pi * j)) ** (2 * p)
                           for j in range(1, GAMMA_TERMS + 1))
            lam = g1 / g2
            bsp[kk] /= lam * lam
    return 1.0 / bsp


def form_factor(sym: str, q) -> np.ndarray:
    """f(q) = sum_l a_l exp(-b_l (q/4pi)^2) + c (opsfact.cu:29-35), float64."""
    v = FORM_FACTORS[sym]
    s2 = (np.asarray(q, dtype=np.float64) / (4.0 * np.pi)) ** 2
    return sum(v[l] * np.exp(-v[4 + l] * s2) for l in range(4)) + v[8]


# ----------------------------------------------------------------- plan ---

@dataclass
class OraclePlan:
    n: Tuple[int, int, int]
    nn: Tuple[int, int, int]
    order: int
    hp: HistParams
    moduli: Tuple[np.ndarray, np.ndarray, np.ndarray]
    border: int
    # None: bulk-solvent pad = border-slab mean (padding::avg, the
    # reference's only live mode).  A dict {symbol: atoms per cell voxel}:
    # padding::given, elements not in the dict padded with 0.
    pad_given: Optional[Dict[str, float]] = None
    # fractional-coordinate matrix u = M x and reciprocal rows b_a (q = sum m_a b_a)
    M: Optional[np.ndarray] = None
    basis: Optional[np.ndarray] = None
    _bin_cache: Dict[tuple, int] = field(default_factory=dict)

    @property
    def triclinic(self) -> bool:
        return self.M is not None and bool(np.any(self.M != np.diag(np.diag(self.M))))

    @property
    def isotropic(self) -> bool:
        s = self.hp.spacing
        return bool(s[0] == s[1] == s[2]) and not self.triclinic

    def bins_for(self, m: np.ndarray) -> np.ndarray:
        """Bins of frequencies given as rows (m_x, m_y, m_z): by the shell
        n = |m|^2 when the spacings are equal, by (|m_x|, |m_y|, |m_z|) for other
        orthorhombic grids, by the signed m for triclinic cells."""
        iso, tric = self.isotropic, self.triclinic
        if iso:
            keys = (m ** 2).sum(axis=1)[:, None]
        else:
            keys = m if tric else np.abs(m)
        uniq, inv = np.unique(keys, axis=0, return_inverse=True)
        vals = np.empty(len(uniq), dtype=np.int64)
        hp = self.hp
        for i, row in enumerate(uniq):
            k = tuple(int(v) for v in row)
            if k not in self._bin_cache:
                if iso:
                    b = bin_of_n(k[0], hp.spacing[0], hp.dq, hp.qcut, hp.num_bins)
                elif tric:
                    b = bin_of_vec(k, self.basis, hp.dq, hp.qcut, hp.num_bins)
                else:
                    b = bin_of_m(k, hp.spacing, hp.dq, hp.qcut, hp.num_bins)
                self._bin_cache[k] = b
            vals[i] = self._bin_cache[k]
        return vals[np.asarray(inv).reshape(-1)]


def make_plan(n, box, order=4, sigma=2.0, scaled=None, dq_req=0.0, qcut_req=0.0,
              dq_policy="reference", pad_given=None) -> OraclePlan:
    n = tuple(int(x) for x in n)
    nn = tuple(int(x) for x in scaled) if scaled else tuple(find_closest_product(x, sigma) for x in n)
    M = inverse_box_matrix(box)
    hp = hist_params(n, nn, np.diag(M), sigma, dq_req, qcut_req, dq_policy)
    mods = tuple(bsp_moduli(nn[a], order) for a in range(3))
    pads = None if pad_given is None else {str(k): float(v) for k, v in dict(pad_given).items()}
    return OraclePlan(n, nn, order, hp, mods, border_planes(n[2]), pads, M,
                      reciprocal_basis(M, n, nn))


# Water densities in g/cm^3 and the constants of AtomCounter.cpp:20-29.
WATER_DENSITY = {"SPCE": 0.998, "SPC": 0.998, "TIP3P": 0.982, "TIP4P": 0.999}
AVOGADRO = 6.022e23
WATER_MOLAR_MASS = 18.015


def solvent_pads(lengths, n, water_model: str, ions: Dict[str, int]) -> Dict[str, float]:
    """Atoms per cell voxel of the explicit solvent, AtomCounter::
    calculateAtomCounts (AtomCounter.cpp:56-103) in its float32 arithmetic:
    n_w = rho_w(model) * V * 1e-24 / 18.015 * 6.022e23 water molecules in
    the box, O = n_w / (n_x n_y n_z), H = 2 n_w / (n_x n_y n_z), each ion
    count / (n_x n_y n_z).  An unknown model raises, as std::map::at does
    there (AtomCounter.cpp:58 looks the model up before the SPCE fallback)."""
    if water_model not in WATER_DENSITY:
        raise ValueError(f"unknown water model {water_model!r}")
    lx, ly, lz = (F32(v) for v in lengths)
    volume = F32(F32(lx * ly) * lz)
    volume_cm3 = F32(volume * F32(1e-24))
    mass = F32(F32(WATER_DENSITY[water_model]) * volume_cm3)
    molecules = F32(F32(mass / F32(WATER_MOLAR_MASS)) * F32(AVOGADRO))
    total = int(n[0]) * int(n[1]) * int(n[2])
    tot = F32(total)
    out = {"O": float(F32(molecules / tot)), "H": float(F32(F32(2.0) * molecules / tot))}
    for ion, count in ions.items():
        out[str(ion)] = float(F32(F32(int(count)) / tot))
    return out


# ------------------------------------------------------------ per frame ---

# When True, grid coordinates are computed in float64 from float64
# positions (used only to pin the adjoint against finite differences, which
# float32 coordinate rounding would blur); the default follows the reference.
EXACT_COORDS = False


def grid_coords(pos, oc, n):
    """Cell index m and fraction w per atom and axis, following the
    reference's float32/float64 mix (saxsDeviceKernels.cu:352-368).  `oc` is
    the diagonal of an orthorhombic inverse box, or the full matrix M; for a
    non-diagonal M, u_a = (M_a0 x + M_a1 y) + M_a2 z with every product and
    sum rounded to float32 (no fused multiply-add)."""
    oc = np.asarray(oc, dtype=np.float32)
    if oc.ndim == 2 and not np.any(oc != np.diag(np.diag(oc))):
        oc = np.diag(oc).copy()
    if EXACT_COORDS:
        pos = np.asarray(pos, dtype=np.float64)
        if oc.ndim == 2:
            ud = pos @ oc.astype(np.float64).T
        else:
            ud = pos * oc.astype(np.float64)[None, :]
        r = np.asarray(n, dtype=np.float64)[None, :] * (ud - np.rint(ud - 0.5))
        m = np.floor(r).astype(np.int64)
        return m, r - m
    pos = np.asarray(pos, dtype=np.float32)
    if oc.ndim == 2:
        u = np.empty_like(pos)
        for a in range(3):
            t = (oc[a, 0] * pos[:, 0]).astype(np.float32) + (oc[a, 1] * pos[:, 1]).astype(np.float32)
            u[:, a] = t.astype(np.float32) + (oc[a, 2] * pos[:, 2]).astype(np.float32)
    else:
        u = (pos * oc[None, :]).astype(np.float32)
    ud = u.astype(np.float64)
    r = (np.asarray(n, dtype=np.float64)[None, :] * (ud - np.rint(ud - 0.5))).astype(np.float32)
    m = r.astype(np.int64)                      # truncation of non-negative floats
    w = (r - m.astype(np.float32)).astype(np.float32)
    return m, w


def spread(pos, oc_diag, n, p) -> np.ndarray:
    """Unit-weight B-spline density of one element (S1), float64 grid.
    Grid point (m_a - p + o) (negative indices wrapped by +n_a) receives
    theta_x[o] theta_y[o'] theta_z[o'']."""
    nx, ny, nz = n
    grid = np.zeros(nx * ny * nz, dtype=np.float64)
    if len(pos) == 0:
        return grid.reshape(n)
    m, w = grid_coords(pos, oc_diag, n)
    th = [bspline_theta(np.asarray(w[:, a], dtype=np.float64), p) for a in range(3)]
    idx = []
    for a in range(3):
        i = m[:, a][:, None] - p + np.arange(p)[None, :]
        i = np.where(i < 0, i + n[a], i)
        idx.append(i)
    flat = (idx[0][:, :, None, None] * ny + idx[1][:, None, :, None]) * nz + idx[2][:, None, None, :]
    wts = th[0][:, :, None, None] * th[1][:, None, :, None] * th[2][:, None, None, :]
    grid += np.bincount(flat.ravel(), weights=wts.ravel(), minlength=grid.size)
    return grid.reshape(n)


def border_pad(grid, d) -> float:
    """Mean density over the z-slabs z <= d or z >= nz-d (S2)."""
    nz = grid.shape[2]
    mask = np.zeros(nz, dtype=bool)
    mask[: d + 1] = True
    mask[nz - d:] = True
    sel = grid[:, :, mask]
    return float(sel.sum()) / sel.size


def half_spectrum_m(nn):
    """Signed frequencies of the R2C half spectrum (reference convention
    m = i < nf ? i : i - nn)."""
    out = []
    for a, N in enumerate(nn):
        nf = N // 2 if N % 2 == 0 else N // 2 + 1
        L = N if a < 2 else N // 2 + 1
        i = np.arange(L)
        out.append(np.where(i < nf, i, i - N))
    return out


def spectrum_bins(plan: OraclePlan):
    """Bin of every half-spectrum point (-1 outside q <= qcut or from
    num_bins on), the in-range mask and |q|, for the plan's box."""
    mx, my, mz = half_spectrum_m(plan.nn)
    bv = plan.basis.astype(np.float64)
    qv = (mx[:, None, None, None] * bv[0] + my[None, :, None, None] * bv[1]
          + mz[None, None, :, None] * bv[2])
    q2 = (qv ** 2).sum(axis=3)
    bins = np.full(q2.shape, -1, dtype=np.int64)
    cand = q2 <= float(plan.hp.qcut) ** 2 * (1.0 + 1e-9)
    ii, jj, kk = np.nonzero(cand)
    bins[cand] = plan.bins_for(np.stack([mx[ii], my[jj], mz[kk]], axis=1))
    return bins, bins >= 0, np.sqrt(q2)


def frame_spectrum_terms(plan: OraclePlan, pos, elem_ids, symbols, box, literal=True):
    """Per-element F_e over the half spectrum plus the in-range mask and q.

    literal=True embeds exactly like the reference (pad everywhere outside
    the cell, S3); literal=False transforms (rho - pad) inside the cell with
    zero padding, which differs only at m = 0."""
    oc = inverse_box_matrix(box)
    bins, inr, q = spectrum_bins(plan)
    Fs = []
    for e, sym in enumerate(symbols):
        sel = np.asarray(elem_ids) == e
        g = spread(np.asarray(pos)[sel], oc, plan.n, plan.order)
        pad = border_pad(g, plan.border) if plan.pad_given is None \
            else plan.pad_given.get(sym, 0.0)
        if literal:
            sup = np.full(plan.nn, pad, dtype=np.float64)
            sup[: plan.n[0], : plan.n[1], : plan.n[2]] = g
        else:
            sup = np.zeros(plan.nn, dtype=np.float64)
            sup[: plan.n[0], : plan.n[1], : plan.n[2]] = g - pad
        Fs.append(_rfftn(sup))
    return Fs, inr, bins, q


def frame_intensity(plan, pos, elem_ids, symbols, box, literal=True):
    """I = |sum_e f_e(q) F_e|^2 * modX modY modZ over the half spectrum
    (S4 + S5); zero outside q <= qcut like the reference's accumulator."""
    Fs, inr, bins, q = frame_spectrum_terms(plan, pos, elem_ids, symbols, box, literal)
    A = np.zeros(Fs[0].shape if Fs else inr.shape, dtype=np.complex128)
    for e, sym in enumerate(symbols):
        A[inr] += form_factor(sym, q[inr]) * Fs[e][inr]
    mods = plan.moduli
    I = (A.real ** 2 + A.imag ** 2) * mods[0][:, None, None] * mods[1][None, :, None] \
        * mods[2][None, None, : A.shape[2]]
    return I, inr, bins


def histogram(plan, I, inr, bins):
    """S6, literally: every in-range half-spectrum point adds I[i,j,k] +
    I[-i,-j,k] with count 2; the origin adds I once with count 1.  In a
    triclinic cell (-i,-j,k) is not at the same |q| as (i,j,k), so each point
    adds 2 I at its own bin instead (I(-m) = I(m) at q(-m) = -q(m): the point
    stands for itself and its Hermitian mirror)."""
    nb = plan.hp.num_bins
    H = np.zeros(nb)
    C = np.zeros(nb)
    nnx, nny = plan.nn[0], plan.nn[1]
    ii, jj, kk = np.nonzero(inr)
    ib = (nnx - ii) % nnx
    jb = (nny - jj) % nny
    b = bins[ii, jj, kk]
    origin = (ii == 0) & (jj == 0) & (kk == 0)
    mirror = 2.0 * I[ii, jj, kk] if plan.triclinic else I[ii, jj, kk] + I[ib, jb, kk]
    v = np.where(origin, I[ii, jj, kk], mirror)
    c = np.where(origin, 1.0, 2.0)
    np.add.at(H, b, v)
    np.add.at(C, b, c)
    return H, C


def frame_histogram(plan, pos, elem_ids, symbols, box, literal=True):
    I, inr, bins = frame_intensity(plan, pos, elem_ids, symbols, box, literal)
    return histogram(plan, I, inr, bins)


def saxs_rows(plan, H, C):
    """Output rows b = 1..num_bins-1 with C_b != 0: (b*dq, H_b/C_b)
    (saxsKernel.cu:216-223).  q is the float32 product b*dq."""
    rows = []
    for b in range(1, plan.hp.num_bins):
        if C[b] != 0.0:
            rows.append((float(F32(F32(b) * plan.hp.dq)), H[b] / C[b]))
    return rows


def format_rows(rows) -> str:
    """'%10.5f%12.5e' per line (RunSaxs.cu:224-227)."""
    return "".join("%10.5f%12.5e\n" % (q, v) for q, v in rows)


# ---------------------------------------------------------- multi frame ---

def trajectory_histogram(plan, frames, elem_ids, symbols, literal=True):
    """Accumulate (H, C) over frames (NPT mode, RunSaxs.cu:154-203)."""
    H = np.zeros(plan.hp.num_bins)
    C = np.zeros(plan.hp.num_bins)
    for pos, box in frames:
        h, c = frame_histogram(plan, pos, elem_ids, symbols, box, literal)
        H += h
        C += c
    return H, C


def trajectory_histogram_nvt(plan0, frames, elem_ids, symbols, plan_of_box, literal=True):
    """NVT mode: every frame's intensity, with its own box's q cut and form
    factors (plan_of_box(box)), is summed over the half spectrum
    (gridAddKernel, saxsKernel.cu:184-189), and the sum is binned once with
    the first box's q (plan0; RunSaxs.cu:124-130,208-211), every value times
    1/F (calculate_histogram with frames_fact, saxsKernel.cu:17-39 and
    saxsDeviceKernels.cu:9-66).  Returns (H, C) of that one binning."""
    Iacc = None
    for pos, box in frames:
        I, _, _ = frame_intensity(plan_of_box(box), pos, elem_ids, symbols, box, literal)
        Iacc = I if Iacc is None else Iacc + I
    bins, inr, _ = spectrum_bins(plan0)
    return histogram(plan0, Iacc / float(len(frames)), inr, bins)


# ------------------------------------------------------------- adjoint ---

def binned_intensity(plan, pos, elem_ids, symbols, box):
    """Mean intensity per bin, I_b = H_b / C_b (0 where C_b = 0)."""
    H, C = frame_histogram(plan, pos, elem_ids, symbols, box, literal=False)
    return np.where(C > 0, H / np.where(C > 0, C, 1.0), 0.0), C


def grad_binned_intensity(plan, pos, elem_ids, symbols, box, gbar):
    """Gradient of L = sum_b gbar_b * I_b with I_b = H_b/C_b."""
    H, C = frame_histogram(plan, pos, elem_ids, symbols, box, literal=False)
    gH = np.where(C > 0, np.asarray(gbar, dtype=np.float64) / np.where(C > 0, C, 1.0), 0.0)
    return grad_linear_H(plan, pos, elem_ids, symbols, box, gH)


def grad_linear_H(plan, pos, elem_ids, symbols, box, gH):
    """Analytic gradient of L = sum_b gH_b * H_b with respect to atom
    positions, float64, by the transposed pipeline: histogram^T -> |.|^2 ->
    form factor -> R2C^T (half-spectrum inverse DFT, real part) -> crop +
    pad-mean term -> B-spline gradient gather."""
    pos = np.asarray(pos, dtype=np.float64 if EXACT_COORDS else np.float32)
    oc = inverse_box_matrix(box)
    n, nn, p = plan.n, plan.nn, plan.order
    Fs, inr, bins, q = frame_spectrum_terms(plan, pos, elem_ids, symbols, box, literal=False)
    A = np.zeros(inr.shape, dtype=np.complex128)
    ffs = [form_factor(sym, q) for sym in symbols]
    for e in range(len(symbols)):
        A[inr] += ffs[e][inr] * Fs[e][inr]
    mod = plan.moduli[0][:, None, None] * plan.moduli[1][None, :, None] \
        * plan.moduli[2][None, None, : inr.shape[2]]
    # dL/dI at each half-spectrum point: weight 2 (origin 1) * gH_b
    w = np.where(inr, 2.0, 0.0)
    w[0, 0, 0] = 1.0 if inr[0, 0, 0] else 0.0
    bsafe = np.where(inr, bins, 0)
    dI = w * np.asarray(gH, dtype=np.float64)[bsafe]
    Abar = 2.0 * mod * dI * A                       # dL/d(Re A) + i dL/d(Im A)
    grad = np.zeros((len(pos), 3))
    elem_ids = np.asarray(elem_ids)
    for e in range(len(symbols)):
        Fbar = np.where(inr, ffs[e] * Abar, 0.0)
        # transpose of the half-spectrum map: Re sum_p Fbar(p) e^{+2 pi i p.g/nn}
        full = np.zeros(nn, dtype=np.complex128)
        full[:, :, : Fbar.shape[2]] = Fbar
        rho_t = np.real(_ifftn(full) * (nn[0] * nn[1] * nn[2]))[: n[0], : n[1], : n[2]]
        # pad-mean term: rho~ = rho - pad, pad = mean over border z-slabs
        d = plan.border
        mask = np.zeros(n[2], dtype=bool)
        mask[: d + 1] = True
        mask[n[2] - d:] = True
        count = n[0] * n[1] * int(mask.sum())
        drho = rho_t.copy()
        if plan.pad_given is None:             # a given pad does not move with the atoms
            drho[:, :, mask] -= rho_t.sum() / count
        sel = np.nonzero(elem_ids == e)[0]
        if len(sel) == 0:
            continue
        m, wv = grid_coords(pos[sel], oc, n)
        th = [bspline_theta(np.asarray(wv[:, a], dtype=np.float64), p) for a in range(3)]
        dth = [bspline_dtheta(np.asarray(wv[:, a], dtype=np.float64), p) for a in range(3)]
        idx = []
        for a in range(3):
            i = m[:, a][:, None] - p + np.arange(p)[None, :]
            idx.append(np.where(i < 0, i + n[a], i))
        G = drho[idx[0][:, :, None, None], idx[1][:, None, :, None], idx[2][:, None, None, :]]
        # dL/dr_a (r_a = n_a u_a + const), then x through u = M x
        gr = np.stack([np.einsum("aijk,ai,aj,ak->a", G, dth[0], th[1], th[2]),
                       np.einsum("aijk,ai,aj,ak->a", G, th[0], dth[1], th[2]),
                       np.einsum("aijk,ai,aj,ak->a", G, th[0], th[1], dth[2])], axis=1)
        grad[sel] = (gr * np.asarray(n, dtype=np.float64)[None, :]) @ oc.astype(np.float64)
    return grad


# ------------------------------------------------------------- inputs ---

ELEMENT_MIX_5 = (("H", 0.50), ("C", 0.30), ("O", 0.15), ("N", 0.04), ("S", 0.01))


def element_counts(n_atoms: int, mix=ELEMENT_MIX_5) -> Dict[str, int]:
    counts = {s: int(round(n_atoms * f)) for s, f in mix}
    counts[mix[0][0]] += n_atoms - sum(counts.values())
    return counts


def synthetic_system(n_atoms: int, box_len: float, seed: int, mix=ELEMENT_MIX_5):
    """Seeded uniform atoms with an exact element mix; returns
    (positions float32 [N,3], element symbol per atom)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, box_len, size=(n_atoms, 3)).astype(np.float32)
    syms = []
    for s, c in element_counts(n_atoms, mix).items():
        syms += [s] * c
    syms = list(np.asarray(syms)[rng.permutation(n_atoms)])
    return pos, syms


def frame_positions(base, seed: int, frame: int, sigma_disp: float):
    rng = np.random.default_rng([seed, frame])
    return (base + rng.normal(0.0, sigma_disp, size=base.shape)).astype(np.float32)


def element_ids(syms: Sequence[str]):
    """Symbols in std::map order (byte-wise sorted, RunSaxs.cu:108) and the
    per-atom index into that list."""
    symbols = sorted(set(syms))
    lut = {s: i for i, s in enumerate(symbols)}
    return symbols, np.array([lut[s] for s in syms], dtype=np.int32)
