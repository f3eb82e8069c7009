"""Host-side setup of one system: FFT grid, histogram parameters, shell
tables and B-spline moduli.  These are the per-run computations the
reference does on the host before its frame loop; the results are handed to
the engine's C ABI as plain arrays.

  fft_size             saxsKernel.cu:377-469 (findClosestProduct)
  border_planes        saxsKernel.h:107-110
  inverse_box_diag     Cell.cpp:111-126
  inverse_box_matrix   Cell.cpp:118-125 applied to the transposed (MDAnalysis) box
  hist_params          saxsKernel.cu:489-526 and :247
  shell_bins           exact bin of every shell n = |m|^2
  point_bins           exact bin of every point of an orthorhombic grid
  triclinic_bins       exact bin of every signed point of a triclinic grid
  half_spectrum_counts  the counts calculate_histogram adds (saxsDeviceKernels.cu:67-125)
  bspline_moduli       BSpmod.cpp:45-211
  form_factor_table    opsfact.cu:29-35 on the shell q values
  explicit_solvent_pads  AtomCounter.cpp:56-103 (padding::given, Options.h:7-11)
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Mapping, Optional, Sequence, Tuple

import numpy as np

from .formfactors import coefficients

F32 = np.float32
_SHELL = F32(0.07)            # Parameters.h:8
_GAMMA_TERMS = 50             # BSpmod.h:22
_TINY = 1.0e-7                # BSpmod.cpp:104


def _smooth_numbers(limit: int):
    vals = []
    p2 = 1
    while p2 <= limit:
        p3 = p2
        while p3 <= limit:
            p5 = p3
            while p5 <= limit:
                p7 = p5
                while p7 <= limit:
                    vals.append(p7)
                    p7 *= 7
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return sorted(vals)


def fft_size(n: int, sigma: float) -> int:
    """2,3,5,7-smooth size closest to round(n*sigma) (float32 product, halves
    rounded away from zero); on ties the smaller size wins."""
    t = float(F32(n) * F32(sigma))
    target = int(math.floor(abs(t) + 0.5)) * (1 if t >= 0 else -1)
    best, best_d = target, None
    for v in _smooth_numbers(2 * target):
        d = abs(v - target)
        if best_d is None or d < best_d:
            best, best_d = v, d
    return best


def border_planes(nz: int) -> int:
    return int(F32(F32(_SHELL * F32(nz)) / F32(2.0)))


def max_given_pad(nz: int) -> float:
    """Largest constant pad (atoms per cell voxel) the engine accepts: the
    fixed-point pad (2^32 units per atom) summed over a column of n_z voxels
    stays below 2^63."""
    return 2.0 ** 31 / float(nz)


def inverse_box_diag(box) -> np.ndarray:
    """oc diagonal as float32(1/L_a); raises for a triclinic box."""
    b = np.asarray(box, dtype=np.float64).reshape(3, 3)
    off = b - np.diag(np.diag(b))
    if np.any(off != 0.0):
        raise ValueError("triclinic boxes are not supported: the box must be diagonal "
                         "(orthorhombic)")
    d = np.asarray(np.diag(b), dtype=np.float32)
    if np.any(~np.isfinite(d)) or np.any(d <= 0):
        raise ValueError(f"box lengths must be positive, got {d.tolist()}")
    return np.array([F32(1.0 / float(x)) for x in d], dtype=np.float32)


def inverse_box_matrix(box) -> np.ndarray:
    """Fractional-coordinate matrix M (u = M x, float32, row-major) of a cell
    given MDAnalysis-style: rows a, b, c with a along x and b in the xy plane
    (lower-triangular).  Cartesian x = B^T u; B^T is upper-triangular and
    Cell.cpp:118-125's float32 formula for an upper-triangular inverse gives
    M = (B^T)^-1 exactly as the reference would for that orientation (the
    reference applies it to B and so drops the off-diagonal terms).  A
    diagonal box gives diag(float32(1/L_a)), as inverse_box_diag."""
    b = np.asarray(box, dtype=np.float32).reshape(3, 3)
    if b[0, 1] != 0 or b[0, 2] != 0 or b[1, 2] != 0:
        raise ValueError("box must be lower-triangular: rows a = (ax, 0, 0), b = (bx, by, 0), "
                         "c = (cx, cy, cz)")
    d = np.diag(b)
    if np.any(~np.isfinite(b)) or np.any(d <= 0):
        raise ValueError(f"box diagonal must be positive and finite, got {d.tolist()}")
    t = b.T                                          # upper-triangular, columns a, b, c
    M = np.zeros((3, 3), dtype=np.float32)
    M[0, 0] = F32(1.0 / float(t[0, 0]))
    M[1, 1] = F32(1.0 / float(t[1, 1]))
    M[2, 2] = F32(1.0 / float(t[2, 2]))
    M[0, 1] = F32(-t[0, 1] / F32(t[0, 0] * t[1, 1]))
    M[1, 2] = F32(-t[1, 2] / F32(t[1, 1] * t[2, 2]))
    M[0, 2] = F32(F32(F32(t[0, 1] * t[1, 2]) - F32(t[0, 2] * t[1, 1]))
                  / F32(F32(t[0, 0] * t[1, 1]) * t[2, 2]))
    return M


def inverse_box_matrices(boxes) -> np.ndarray:
    """inverse_box_matrix of a stack of boxes [F, 3, 3], the same float32
    operations elementwise (bit for bit the per-box result)."""
    b = np.asarray(boxes, dtype=np.float32).reshape(-1, 3, 3)
    if np.any(b[:, 0, 1] != 0) or np.any(b[:, 0, 2] != 0) or np.any(b[:, 1, 2] != 0):
        raise ValueError("box must be lower-triangular: rows a = (ax, 0, 0), b = (bx, by, 0), "
                         "c = (cx, cy, cz)")
    d = np.diagonal(b, axis1=1, axis2=2)
    if np.any(~np.isfinite(b)) or np.any(d <= 0):
        raise ValueError("box diagonal must be positive and finite")
    t = np.transpose(b, (0, 2, 1))
    M = np.zeros_like(b)
    one = np.float64(1.0)
    M[:, 0, 0] = (one / t[:, 0, 0].astype(np.float64)).astype(np.float32)
    M[:, 1, 1] = (one / t[:, 1, 1].astype(np.float64)).astype(np.float32)
    M[:, 2, 2] = (one / t[:, 2, 2].astype(np.float64)).astype(np.float32)
    M[:, 0, 1] = -t[:, 0, 1] / (t[:, 0, 0] * t[:, 1, 1])
    M[:, 1, 2] = -t[:, 1, 2] / (t[:, 1, 1] * t[:, 2, 2])
    M[:, 0, 2] = (t[:, 0, 1] * t[:, 1, 2] - t[:, 0, 2] * t[:, 1, 1]) / ((t[:, 0, 0] * t[:, 1, 1])
                                                                       * t[:, 2, 2])
    return M


def reciprocal_bases(Ms, n: Sequence[int], nn: Sequence[int]) -> np.ndarray:
    """reciprocal_basis of a stack of inverse cells [F, 3, 3] (float32,
    the same operations elementwise)."""
    Ms = np.asarray(Ms, dtype=np.float32).reshape(-1, 3, 3)
    tp = F32(F32(2.0) * F32(math.pi))
    ratio = np.array([F32(F32(nn[a]) / F32(n[a])) for a in range(3)], dtype=np.float32)
    return ((tp * Ms) / ratio[None, :, None]).astype(np.float32)


def is_triclinic(M) -> bool:
    M = np.asarray(M)
    return bool(np.any(M != np.diag(np.diag(M))))


def reciprocal_basis(M, n: Sequence[int], nn: Sequence[int]) -> np.ndarray:
    """Rows b_a (float32) with q = sum_a m_a b_a for FFT index m:
    b_a = 2 pi M_a. n_a / nn_a, each entry in the float32 expression of the
    reciprocal spacing (so a diagonal M gives diag(spacing))."""
    tp = F32(F32(2.0) * F32(math.pi))
    out = np.zeros((3, 3), dtype=np.float32)
    for a in range(3):
        ratio = F32(F32(nn[a]) / F32(n[a]))
        for j in range(3):
            out[a, j] = F32(F32(tp * F32(M[a, j])) / ratio)
    return out


def _basis_integers(basis):
    """Integer rows B and a denominator D with basis = B / D exactly."""
    ratios = [[float(v).as_integer_ratio() for v in row] for row in basis]
    D = 1
    for row in ratios:
        for _, den in row:
            D = D * den // math.gcd(D, den)
    return [[num * (D // den) for num, den in row] for row in ratios], D


def triclinic_kmax(basis, qcut: float, nn: Sequence[int]) -> Tuple[int, int, int]:
    """Largest |m_a| of the points with |sum m_a b_a| <= qcut (the ellipsoid's
    extent qcut sqrt((G^-1)_aa), G = b b^T), clipped to nn_a/2; the exact
    table marks the box corners outside qcut."""
    b = np.asarray(basis, dtype=np.float64)
    Ginv = np.linalg.inv(b @ b.T)
    out = []
    for a in range(3):
        k = int(math.floor(float(qcut) * math.sqrt(max(Ginv[a, a], 0.0)) * (1 + 1e-9))) + 1
        out.append(min(k, int(nn[a]) // 2))
    return tuple(out)


def _bins_filtered(q2, err, dq: float, qcut: float, num_bins: int, exact) -> np.ndarray:
    """Bins of points whose squared magnitudes q2 are known in float64 to
    within err: max{b : b^2 dq^2 <= q2}, -1 outside q2 <= qcut^2 or from
    num_bins on.  A point whose q2 lies within 1e-9 q2 + err of qcut^2 or of a
    bin edge b^2 dq^2 is decided by exact(i) in integer arithmetic, so the
    result equals the exact rule everywhere."""
    dq, qcut = float(dq), float(qcut)
    dq2, qc2 = dq * dq, qcut * qcut
    tol = 1e-9 * q2 + err
    t = q2 / dq2
    b = np.floor(np.sqrt(t))
    amb = ((np.abs(q2 - qc2) <= tol) | ((t - b * b) * dq2 <= tol)
           | (((b + 1.0) ** 2 - t) * dq2 <= tol))
    bi = b.astype(np.int64)
    out = np.where((q2 <= qc2) & (bi < num_bins), bi, -1)
    for i in np.flatnonzero(amb):
        out[i] = exact(int(i))
    return out.astype(np.int32)


def triclinic_bins(basis, dq: float, qcut: float, num_bins: int,
                   kmax: Sequence[int]) -> np.ndarray:
    """bin of every signed point m of the box |m_a| <= kmax[a] (layout
    ((m_x+K_x)(2K_y+1) + m_y+K_y)(2K_z+1) + m_z+K_z): q = sum_a m_a b_a with
    the float32 rows as exact integers over one denominator,
    bin = max{b : b^2 dq^2 <= |q|^2}, -1 outside |q| <= qcut or from num_bins.
    Evaluated in float64 with an error bound; points near an edge are
    decided exactly in integers."""
    B, D = _basis_integers(basis)
    dn, dd = float(dq).as_integer_ratio()
    qn, qd = float(qcut).as_integer_ratio()
    kx, ky, kz = (int(v) for v in kmax)
    shape = (2 * kx + 1, 2 * ky + 1, 2 * kz + 1)
    size = shape[0] * shape[1] * shape[2]
    if dn <= 0:
        return np.full(size, -1, dtype=np.int32)
    lim = qn * qn * D * D                          # |q D|^2 qd^2 <= qn^2 D^2
    den = dn * dn * D * D

    def exact(i):
        a, b, c = np.unravel_index(i, shape)
        m = (int(a) - kx, int(b) - ky, int(c) - kz)
        s2 = 0
        for j in range(3):
            v = m[0] * B[0][j] + m[1] * B[1][j] + m[2] * B[2][j]
            s2 += v * v                            # |q|^2 D^2
        if s2 * qd * qd > lim:
            return -1
        bb = math.isqrt((s2 * dd * dd) // den)
        return bb if bb < num_bins else -1

    q2, err = _triclinic_q2(basis, kmax)
    return _bins_filtered(q2, err, dq, qcut, num_bins, exact)


def _quadratic_form(G, kmax, signed: bool = True):
    """sum_ab G_ab m_a m_b over the signed box |m_a| <= kmax[a] (or over
    |m_a| when not signed), flattened in triclinic_bins order."""
    m = [np.arange(-int(k), int(k) + 1, dtype=np.float64) for k in kmax]
    if not signed:
        m = [np.abs(v) for v in m]
    X, Y, Z = m[0][:, None, None], m[1][None, :, None], m[2][None, None, :]
    out = (G[0, 0] * X * X + G[1, 1] * Y * Y) + G[2, 2] * Z * Z
    out = out + (2.0 * G[0, 1] * X * Y + 2.0 * G[0, 2] * X * Z) + 2.0 * G[1, 2] * Y * Z
    return out.reshape(-1)


def _triclinic_q2(basis, kmax):
    """|q|^2 = m^T (b b^T) m over the signed box, and a bound on its float64
    error: 1e-12 sum_ab (|b| |b|^T)_ab |m_a| |m_b|, a thousand times the
    rounding of the Gram matrix and of the form (a few 1e-16 of that sum)."""
    b = np.asarray(basis, dtype=np.float64)
    ab = np.abs(b)
    return _quadratic_form(b @ b.T, kmax), 1e-12 * _quadratic_form(ab @ ab.T, kmax, signed=False)


def form_factor_vectors(symbols: Sequence[str], basis, kmax: Sequence[int],
                        mask: Optional[np.ndarray] = None) -> np.ndarray:
    """f_e(|q|), q = sum_a m_a b_a, over the signed box in triclinic_bins order;
    with `mask`, only at the masked points (0 elsewhere)."""
    q2, _ = _triclinic_q2(basis, kmax)
    q = np.sqrt(np.maximum(q2, 0.0))
    sel = np.ones(len(q), dtype=bool) if mask is None else np.asarray(mask, dtype=bool)
    s2 = (q[sel] / (4.0 * np.pi)) ** 2
    out = np.zeros((len(symbols), len(q)))
    for e, sym in enumerate(symbols):
        a1, a2, a3, a4, b1, b2, b3, b4, c = coefficients(sym)
        out[e, sel] = (a1 * np.exp(-b1 * s2) + a2 * np.exp(-b2 * s2) + a3 * np.exp(-b3 * s2)
                       + a4 * np.exp(-b4 * s2) + c)
    return out


@dataclass(frozen=True)
class Histogram:
    dq: float                 # float32 value
    qcut: float               # float32 value
    num_bins: int
    spacing: Tuple[float, float, float]   # float32 reciprocal spacing per axis
    dq_reset: bool
    qcut_reset: bool


def hist_params(n: Sequence[int], nn: Sequence[int], oc_diag, sigma: float, dq_req: float,
                qcut_req: float, dq_policy: str = "reference") -> Histogram:
    """Bin width and cutoff after the reference's reset: dq is raised to the
    coarsest grid spacing 2*pi*oc/sigma, qcut lowered to the largest grid
    frequency (float32 arithmetic throughout)."""
    tp = F32(F32(2.0) * F32(math.pi))
    sig = F32(sigma)
    arg = [F32(F32(tp * F32(oc_diag[a])) / sig) for a in range(3)]
    nf = [v // 2 if v % 2 == 0 else v // 2 + 1 for v in nn]
    cut = [F32(arg[a] * F32(nf[a] - 1)) for a in range(3)]
    if dq_policy == "reference":
        dq = max(arg + [F32(dq_req)])
    elif dq_policy == "exact":
        if not dq_req > 0:
            raise ValueError("dq_policy='exact' needs a positive bin_size")
        dq = F32(dq_req)
    else:
        raise ValueError(f"unknown dq_policy {dq_policy!r}")
    qcut = min(cut + [F32(qcut_req)])
    num_bins = int(F32(qcut / dq)) + 1
    spacing = tuple(float(F32(F32(tp * F32(oc_diag[a])) / F32(F32(nn[a]) / F32(n[a]))))
                    for a in range(3))
    return Histogram(float(dq), float(qcut), num_bins, spacing,
                     bool(F32(dq) != F32(dq_req)), bool(F32(qcut) != F32(qcut_req)))


def shell_bins(spacing: float, dq: float, qcut: float, num_bins: int) -> np.ndarray:
    """bin[n] for shells n = 0..n_max, where n_max is the largest n with
    n*s^2 <= qcut^2 and bin[n] = max{b : b^2 dq^2 <= n s^2} (or -1 when it
    reaches num_bins).  Exact: the float values are expanded to integer
    ratios and compared in integer arithmetic."""
    sn, sd = float(spacing).as_integer_ratio()
    dn, dd = float(dq).as_integer_ratio()
    qn, qd = float(qcut).as_integer_ratio()
    if sn <= 0 or dn <= 0 or qn < 0:
        return np.full(1, -1 if qn < 0 else 0, dtype=np.int32)
    n_max = (qn * qn * sd * sd) // (sn * sn * qd * qd)
    if n_max > 50_000_000:
        raise ValueError("q range too large for the shell tables")
    num = sn * sn * dd * dd          # n * s^2 / dq^2 = n * num / den
    den = dn * dn * sd * sd
    out = np.empty(n_max + 1, dtype=np.int32)
    for v in range(n_max + 1):
        b = math.isqrt((v * num) // den)
        out[v] = b if b < num_bins else -1
    return out


def point_kmax(spacing: Sequence[float], qcut: float) -> Tuple[int, int, int]:
    """Largest |m_a| with (m_a s_a)^2 <= qcut^2 on each axis (exact)."""
    qn, qd = float(qcut).as_integer_ratio()
    out = []
    for a in range(3):
        sn, sd = float(spacing[a]).as_integer_ratio()
        if sn <= 0 or qn < 0:
            out.append(0)
            continue
        out.append(math.isqrt((qn * qn * sd * sd) // (sn * sn * qd * qd)))
    return tuple(out)


def point_bins(spacing: Sequence[float], dq: float, qcut: float, num_bins: int,
               kmax: Sequence[int]) -> np.ndarray:
    """bin of every point (|m_x|, |m_y|, |m_z|) of the box |m_a| <= kmax[a], laid
    out as (|m_x| (kmax_y+1) + |m_y|) (kmax_z+1) + |m_z|, for an orthorhombic
    reciprocal grid: q^2 = sum_a (m_a s_a)^2, bin = max{b : b^2 dq^2 <= q^2},
    -1 outside q <= qcut or from num_bins on.  Evaluated in float64 (each
    m_a s_a is exact there); points near an edge are decided exactly with the
    float values expanded to integer ratios over one common denominator."""
    sr = [float(v).as_integer_ratio() for v in spacing]
    dn, dd = float(dq).as_integer_ratio()
    qn, qd = float(qcut).as_integer_ratio()
    den = 1
    for _, sd in sr:
        den = den * sd * sd // math.gcd(den, sd * sd)
    w = [sn * sn * (den // (sd * sd)) for sn, sd in sr]     # q^2 * den = sum m_a^2 w_a
    kx, ky, kz = (int(v) for v in kmax)
    shape = (kx + 1, ky + 1, kz + 1)
    if dn <= 0:
        return np.full(shape[0] * shape[1] * shape[2], -1, dtype=np.int32)

    def exact(i):
        ax, ay, az = (int(v) for v in np.unravel_index(i, shape))
        q2 = ax * ax * w[0] + ay * ay * w[1] + az * az * w[2]    # q^2 = q2 / den
        if q2 * qd * qd > qn * qn * den:
            return -1
        b = math.isqrt((q2 * dd * dd) // (dn * dn * den))
        return b if b < num_bins else -1

    ax = [np.arange(k, dtype=np.float64) * float(v) for k, v in zip(shape, spacing)]
    q2 = (ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2
          + ax[2][None, None, :] ** 2).reshape(-1)
    return _bins_filtered(q2, 1e-12 * q2, dq, qcut, num_bins, exact)


def form_factor_points(symbols: Sequence[str], spacing: Sequence[float],
                       kmax: Sequence[int]) -> np.ndarray:
    """f_e(q) at q = sqrt(sum_a (m_a s_a)^2) for every point of the box
    |m_a| <= kmax[a] in point_bins order (float64)."""
    ax = [np.arange(int(k) + 1, dtype=np.float64) * float(s) for k, s in zip(kmax, spacing)]
    q2 = (ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2)
    q = np.sqrt(q2.reshape(-1))
    s2 = (q / (4.0 * np.pi)) ** 2
    out = np.empty((len(symbols), len(q)))
    for e, sym in enumerate(symbols):
        a1, a2, a3, a4, b1, b2, b3, b4, c = coefficients(sym)
        out[e] = (a1 * np.exp(-b1 * s2) + a2 * np.exp(-b2 * s2) + a3 * np.exp(-b3 * s2)
                  + a4 * np.exp(-b4 * s2) + c)
    return out


def half_spectrum_m(nn: Sequence[int]):
    """Signed frequencies of the R2C half spectrum per axis: index i maps to
    i below nf = nn/2 (even) or nn/2 + 1 (odd), else i - nn; z runs over
    0..nn_z/2 (saxsDeviceKernels.cu:75-83)."""
    out = []
    for a, N in enumerate(nn):
        N = int(N)
        i = np.arange(N if a < 2 else N // 2 + 1, dtype=np.int64)
        nf = N // 2 if N % 2 == 0 else N // 2 + 1
        out.append(np.where(i < nf, i, i - N))
    return out


def half_spectrum_counts(nn: Sequence[int], bins: np.ndarray, kmax, num_bins: int,
                         signed: bool = False) -> np.ndarray:
    """C_b of one binning of the half spectrum with the tables (bins, kmax)
    of the engine's q tables: every point with a bin adds 2, the origin 1
    (saxsDeviceKernels.cu:100-121).  kmax None means shell keys n = |m|^2;
    otherwise point keys (|m_x|, |m_y|, |m_z|), or signed m when `signed`."""
    bins = np.asarray(bins, dtype=np.int64)
    if kmax is None:
        K = (math.isqrt(len(bins) - 1),) * 3
    else:
        K = tuple(int(k) for k in kmax)
    mx, my, mz = (m[np.abs(m) <= k] for m, k in zip(half_spectrum_m(nn), K))
    X, Y, Z = mx[:, None, None], my[None, :, None], mz[None, None, :]
    if kmax is None:
        key = X * X + Y * Y + Z * Z
        b = np.where(key < len(bins), bins[np.minimum(key, len(bins) - 1)], -1)
    elif signed:
        b = bins[((X + K[0]) * (2 * K[1] + 1) + Y + K[1]) * (2 * K[2] + 1) + Z + K[2]]
    else:
        b = bins[(np.abs(X) * (K[1] + 1) + np.abs(Y)) * (K[2] + 1) + np.abs(Z)]
    w = np.full(b.shape, 2.0)
    w[(X == 0) & (Y == 0) & (Z == 0)] = 1.0
    ok = b >= 0
    return np.bincount(b[ok], weights=w[ok], minlength=num_bins)[:num_bins]


def bspline_moduli(N: int, order: int) -> np.ndarray:
    """lambda(k)^2 / |b(k)|^2 for k = 0..N-1, float64."""
    w0 = _bspline_at_zero(order)
    A = np.zeros(order + 1)
    A[1:] = w0
    k = np.arange(N, dtype=np.float64)[:, None]
    t = np.arange(order + 1, dtype=np.float64)[None, :]
    ang = 2.0 * np.pi * k * t / N
    b2 = (A * np.cos(ang)).sum(1) ** 2 + (A * np.sin(ang)).sum(1) ** 2
    for i in range(1, N - 1):
        if not b2[i] > _TINY:
            b2[i] = 0.5 * (b2[i - 1] + b2[i + 1])
    m = np.arange(N)
    m = np.where(m < N // 2, m, m - N).astype(np.float64)
    x = np.pi * m / N
    j = np.arange(1, _GAMMA_TERMS + 1, dtype=np.float64)[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        xp = x[:, None] / (x[:, None] + np.pi * j)
        xm = x[:, None] / (x[:, None] - np.pi * j)
        g1 = 1.0 + (xp ** order + xm ** order).sum(1)
        g2 = 1.0 + (xp ** (2 * order) + xm ** (2 * order)).sum(1)
        lam = np.where(m == 0, 1.0, g1 / g2)
    return lam * lam / b2


def _bspline_at_zero(order: int) -> np.ndarray:
    """M_p(p-1), ..., M_p(0): the weights of an atom on a grid node."""
    th = np.zeros(order)
    th[0], th[1] = 1.0, 0.0
    for k in range(3, order + 1):
        div = 1.0 / (k - 1)
        th[k - 1] = 0.0
        for j in range(1, k - 1):
            th[k - j - 1] = div * (j * th[k - j - 2] + (k - j) * th[k - j - 1])
        th[0] = div * th[0]
    return th


def form_factor_table(symbols: Sequence[str], spacing: float, n_shells: int) -> np.ndarray:
    """f_e(q_n), q_n = s*sqrt(n), for every element and shell (float64)."""
    q = float(spacing) * np.sqrt(np.arange(n_shells, dtype=np.float64))
    s2 = (q / (4.0 * np.pi)) ** 2
    out = np.empty((len(symbols), n_shells))
    for e, sym in enumerate(symbols):
        a1, a2, a3, a4, b1, b2, b3, b4, c = coefficients(sym)
        out[e] = (a1 * np.exp(-b1 * s2) + a2 * np.exp(-b2 * s2) + a3 * np.exp(-b3 * s2)
                  + a4 * np.exp(-b4 * s2) + c)
    return out


# g/cm^3 per water model, Avogadro's number and the molar mass of water
# (AtomCounter.cpp:20-29).
WATER_MODELS = {"SPCE": 0.998, "SPC": 0.998, "TIP3P": 0.982, "TIP4P": 0.999}
_N_A = F32(6.022e23)
_M_WATER = F32(18.015)


def explicit_solvent_pads(lengths: Sequence[float], n: Sequence[int], water_model: str,
                          ions: Mapping[str, int]) -> Dict[str, float]:
    """Per-voxel atom counts of a box of explicit water plus ions: the values
    of padding::given.  AtomCounter::calculateAtomCounts evaluates, in float32,
    water molecules = (rho_w * L_x L_y L_z * 1e-24 / M_w) * N_A, then
    O = molecules / cells, H = 2 molecules / cells and ion = count / cells
    over the n_x n_y n_z cell grid (AtomCounter.cpp:56-103).  That code looks
    the model up with std::map::at before its SPCE fallback, so an unknown
    model is an error there too."""
    if water_model not in WATER_MODELS:
        raise ValueError(f"Unknown water model '{water_model}' "
                         f"(expected one of {', '.join(sorted(WATER_MODELS))})")
    v = F32(1.0)
    for a in range(3):
        v = F32(v * F32(lengths[a]))
    grams = F32(F32(WATER_MODELS[water_model]) * F32(v * F32(1e-24)))
    molecules = F32(F32(grams / _M_WATER) * _N_A)
    cells = F32(int(n[0]) * int(n[1]) * int(n[2]))
    pads = {"O": float(F32(molecules / cells)), "H": float(F32(F32(F32(2.0) * molecules) / cells))}
    for ion, count in ions.items():
        pads[str(ion)] = float(F32(F32(int(count)) / cells))
    return pads
