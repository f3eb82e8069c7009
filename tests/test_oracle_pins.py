"""Pins the FP64 oracle against facts the mathematics fixes independently of
its own formulas: closed forms, brute-force sums, known shell counts."""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import saxs_oracle as O


def cardinal_bspline(u, p):
    """M_p(u) by the truncated-power closed form
    M_p(u) = 1/(p-1)! sum_k (-1)^k C(p,k) (u-k)_+^(p-1)."""
    s = 0.0
    for k in range(p + 1):
        t = u - k
        if t > 0:
            s += (-1) ** k * math.comb(p, k) * t ** (p - 1)
    return s / math.factorial(p - 1)


@pytest.mark.parametrize("p", [4, 6])
def test_bspline_weights_closed_form(p):
    for w in np.linspace(0.0, 1.0, 17):
        th = O.bspline_theta(w, p)
        ref = [cardinal_bspline(w + p - 1 - k, p) for k in range(p)]
        assert np.allclose(th, ref, atol=1e-14, rtol=0)
        assert abs(th.sum() - 1.0) < 1e-14             # partition of unity


@pytest.mark.parametrize("p", [4, 6])
def test_bspline_derivative_identity(p):
    h = 1e-6
    for w in np.linspace(0.05, 0.95, 9):
        d = O.bspline_dtheta(w, p)
        fd = (O.bspline_theta(w + h, p) - O.bspline_theta(w - h, p)) / (2 * h)
        assert np.allclose(d, fd, atol=1e-8)
        assert abs(d.sum()) < 1e-14


def test_moduli_closed_form_order4():
    N = 128
    k = np.arange(N)
    b2 = (2.0 / 3.0 + np.cos(2 * np.pi * k / N) / 3.0) ** 2
    mod = O.bsp_moduli(N, 4)
    lam2 = mod * b2
    # lambda(0) = 1; lambda from the finite series G_4/G_8 evaluated with mpmath-free
    # direct sums at a few points
    assert abs(lam2[0] - 1.0) < 1e-15
    for kk in (1, 7, 40, 100):
        m = kk if kk < N // 2 else kk - N
        x = math.pi * m / N
        g = lambda r: 1 + sum((x / (x + math.pi * j)) ** r + (x / (x - math.pi * j)) ** r
                              for j in range(1, 51))
        lam = g(4) / g(8)
        assert abs(lam2[kk] - lam * lam) < 1e-13


def test_moduli_closed_form_order6():
    N = 64
    k = np.arange(N)
    t = 2 * np.pi * k / N
    b2 = ((66 + 52 * np.cos(t) + 2 * np.cos(2 * t)) / 120.0) ** 2
    mod = O.bsp_moduli(N, 6)
    assert abs(mod[0] * b2[0] - 1.0) < 1e-15
    # mod * |b|^2 = lambda^2 with lambda = G_6/G_12 >= 1 (even orders)
    lam2 = mod * b2
    assert np.all(lam2 >= 1.0 - 1e-15)
    for kk in (1, 9, 31, 50):
        m = kk if kk < N // 2 else kk - N
        x = math.pi * m / N
        g = lambda r: 1 + sum((x / (x + math.pi * j)) ** r + (x / (x - math.pi * j)) ** r
                              for j in range(1, 51))
        assert abs(lam2[kk] - (g(6) / g(12)) ** 2) < 1e-13


def test_rfftn_matches_bruteforce_dft():
    rng = np.random.default_rng(0)
    n = 6
    x = rng.normal(size=(n, n, n))
    F = np.fft.rfftn(x)
    g = np.arange(n)
    for (a, b, c) in [(0, 0, 0), (1, 2, 3), (5, 1, 2), (3, 3, 0)]:
        ph = np.exp(-2j * np.pi * (a * g[:, None, None] + b * g[None, :, None]
                                   + c * g[None, None, :]) / n)
        assert abs((x * ph).sum() - F[a, b, c]) < 1e-10


def test_spread_single_atom_stencil_and_mass():
    n = (12, 12, 12)
    oc = np.array([1 / 24.0] * 3, dtype=np.float32)
    pos = np.array([[5.3, 13.9, 0.4]], dtype=np.float32)
    g = O.spread(pos, oc, n, 4)
    assert abs(g.sum() - 1.0) < 1e-14
    m, w = O.grid_coords(pos, oc, n)
    th = [O.bspline_theta(float(w[0, a]), 4) for a in range(3)]
    for o, p2, q in itertools.product(range(4), repeat=3):
        idx = [(m[0, a] - 4 + (o, p2, q)[a]) % n[a] for a in range(3)]
        assert abs(g[tuple(idx)] - th[0][o] * th[1][p2] * th[2][q]) < 1e-15
    assert np.count_nonzero(g) == 64


def test_spread_wraps_only_negative_indices():
    # an atom at u just below 1 puts m = n, stencil n-4..n-1, no wrap
    n = (16, 16, 16)
    oc = np.array([1 / 16.0] * 3, dtype=np.float32)
    pos = np.array([[15.99, 0.01, 8.0]], dtype=np.float32)
    m, w = O.grid_coords(pos, oc, n)
    assert m[0, 0] == 15 and m[0, 1] == 0
    g = O.spread(pos, oc, n, 4)
    assert g[12:16, :, :].sum() > 0.999


def test_on_node_atoms_direct_structure_factor():
    """Atoms on grid nodes with pad 0 (no density in the border slabs): the
    pipeline must reproduce |sum_j f_j e^{-2 pi i m.g_j/nn}|^2 prod lambda^2."""
    L = 32.0
    n = 16
    box = np.diag([L] * 3)
    plan = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, 100.0)
    d = plan.border                          # z <= d or z >= n-d is border
    rng = np.random.default_rng(5)
    # nodes whose stencils (g-4 .. g-1) neither wrap around the cell (x, y >= 4)
    # nor reach the border slabs (z >= d + 5), so pad = 0 and the spread is a
    # pure translation of theta(0) per atom
    zs = rng.integers(d + 5, n - d, size=6)
    nodes = np.stack([rng.integers(4, n, 6), rng.integers(4, n, 6), zs], axis=1)
    pos = (nodes * (L / n)).astype(np.float32)
    syms = ["C", "O", "C", "N", "O", "C"]
    symbols, ids = O.element_ids(syms)
    I, inr, bins = O.frame_intensity(plan, pos, ids, symbols, box)
    nn = plan.nn[0]
    mx, my, mz = O.half_spectrum_m(plan.nn)
    # on a node the stencil is theta(0) = (1/6, 4/6, 1/6, 0) at offsets -4..-1,
    # i.e. the atom's weight sits at g-3, g-2, g-1: the grid-node structure
    # factor times b(m)^* per axis; |.|^2 * mod gives lambda^2 exactly
    ii, jj, kk = np.nonzero(inr)
    for t in range(0, len(ii), max(1, len(ii) // 60)):
        i, j, k = ii[t], jj[t], kk[t]
        m = np.array([mx[i], my[j], mz[k]])
        q = float(plan.hp.spacing[0]) * math.sqrt(float(m @ m))
        A = 0j
        for a, s in zip(nodes, syms):
            A += O.form_factor(s, q) * np.exp(-2j * np.pi * (m @ a) / nn)
        lam2 = 1.0
        for ax in range(3):
            x = math.pi * m[ax] / nn
            if m[ax] != 0:
                g = lambda r: 1 + sum((x / (x + math.pi * jj_)) ** r + (x / (x - math.pi * jj_)) ** r
                                      for jj_ in range(1, 51))
                lam2 *= (g(4) / g(8)) ** 2
        expect = abs(A) ** 2 * lam2
        assert abs(I[i, j, k] - expect) <= 1e-9 * max(1.0, expect), (m, I[i, j, k], expect)


def _count_reference_loop(n_cell, nn, L, qcut, dq):
    """Bin counts by enumerating the reference's histogram loop over the half
    spectrum with exact arithmetic (counts 2 per point, 1 at the origin)."""
    plan = O.make_plan((n_cell,) * 3, np.diag([L] * 3), 4, 2.0, None, dq, qcut)
    s = Fraction(float(plan.hp.spacing[0]))
    D = Fraction(float(plan.hp.dq))
    Q = Fraction(float(plan.hp.qcut))
    C = [0] * plan.hp.num_bins
    M = int(float(Q / s)) + 1
    for i in range(-M, M + 1):
        for j in range(-M, M + 1):
            for k in range(0, M + 1):
                n = i * i + j * j + k * k
                q2 = n * s * s
                if q2 > Q * Q:
                    continue
                b = 0
                while (b + 1) ** 2 * D * D <= q2:
                    b += 1
                if b >= plan.hp.num_bins:
                    continue
                C[b] += 1 if n == 0 else 2
    return C, plan


def test_known_shell_counts_64_grid_37A():
    C, plan = _count_reference_loop(64, 128, 37.0, 0.5, 0.01)
    assert plan.nn == (128, 128, 128)
    assert plan.hp.num_bins == 6
    assert C[1:] == [34, 82, 178, 258, 402]


def test_known_shell_counts_128_grid_100A():
    C, plan = _count_reference_loop(128, 256, 100.0, 0.6, 0.005)
    assert plan.hp.num_bins == 20
    assert C[1:] == [34, 82, 178, 258, 450, 506, 786, 922, 1226, 1426, 1690, 2050, 2458,
                     2626, 3170, 3402, 4042, 4362, 506]


def test_shell_counts_are_sums_of_squares():
    """C_b = sum_{n in b} r3(n) + r2(n) (the k_z = 0 plane counted twice)."""
    def r(dim, n):
        M = math.isqrt(n)
        return sum(1 for v in itertools.product(range(-M, M + 1), repeat=dim)
                   if sum(x * x for x in v) == n)
    C, plan = _count_reference_loop(64, 128, 37.0, 0.5, 0.01)
    for b in range(1, 6):
        # shells of bin b, cut at q <= qcut (n s^2 <= 0.25: n <= 34 here)
        ns = [n for n in range(0, 40) if math.isqrt(n) == b and n <= 34]
        assert C[b] == sum(r(3, n) + r(2, n) for n in ns)


def test_bin_rule_matches_oracle_histogram_counts():
    pos, syms = O.synthetic_system(500, 37.0, seed=1)
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((64,) * 3, np.diag([37.0] * 3), 4, 2.0, None, 0.01, 0.5)
    H, C = O.frame_histogram(plan, pos, ids, symbols, np.diag([37.0] * 3))
    assert list(C[1:].astype(int)) == [34, 82, 178, 258, 402]


def test_pad_subtraction_equivalence():
    """Literal pad embedding and (rho - pad) zero padding agree for m != 0."""
    pos, syms = O.synthetic_system(800, 24.0, seed=2)
    symbols, ids = O.element_ids(syms)
    box = np.diag([24.0] * 3)
    plan = O.make_plan((16,) * 3, box, 4, 2.0, None, 0.0, 100.0)
    H1, C1 = O.frame_histogram(plan, pos, ids, symbols, box, literal=True)
    H2, C2 = O.frame_histogram(plan, pos, ids, symbols, box, literal=False)
    assert np.array_equal(C1, C2)
    assert np.allclose(H1[1:], H2[1:], rtol=1e-12)


def test_find_closest_product():
    assert O.find_closest_product(64, 2.0) == 128
    assert O.find_closest_product(128, 2.5) == 320
    assert O.find_closest_product(37, 2.0) == 75          # 74 -> 75 (|1|) beats 72 (|2|)
    assert O.find_closest_product(13, 1.0) == 12          # tie 12/14 -> smaller
    for v in O.generate_multiples(500):
        m = v
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        assert m == 1


def test_border_planes():
    assert [O.border_planes(n) for n in (64, 128, 256)] == [2, 4, 8]


def test_given_pad_without_atoms_is_dirichlet_kernel():
    """padding::given with no atoms: F_e = pad_e * (nn^3 delta - prod_a D(m_a))
    with |D(m)| = |sin(pi m n/nn) / sin(pi m/nn)| (D(0) = n), so
    I = (sum_e f_e pad_e)^2 prod_a D(m_a)^2 * modulus away from the origin;
    zero at even m != 0 on every axis for nn = 2n."""
    n, nn, L = 8, 16, 12.0
    box = np.diag([L] * 3)
    pads = {"O": 0.031, "H": 0.062}
    plan = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, 100.0, pad_given=pads)
    symbols = ["H", "O"]
    empty = np.zeros((0, 3))
    for literal in (True, False):
        I, inr, bins = O.frame_intensity(plan, empty, np.zeros(0, dtype=np.int64), symbols, box,
                                         literal)
        mx, my, mz = O.half_spectrum_m(plan.nn)

        def dirichlet(m):
            m = np.asarray(m, dtype=np.float64)
            out = np.full(m.shape, float(n))
            nz = m != 0
            out[nz] = np.abs(np.sin(np.pi * m[nz] * n / nn) / np.sin(np.pi * m[nz] / nn))
            return out

        D = dirichlet(mx)[:, None, None] * dirichlet(my)[None, :, None] * dirichlet(mz)[None, None, :]
        nsq = mx[:, None, None] ** 2 + my[None, :, None] ** 2 + mz[None, None, :] ** 2
        q = float(plan.hp.spacing[0]) * np.sqrt(nsq.astype(np.float64))
        amp = sum(O.form_factor(s, q) * pads[s] for s in symbols)
        mod = plan.moduli[0][:, None, None] * plan.moduli[1][None, :, None] \
            * plan.moduli[2][None, None, : I.shape[2]]
        expect = amp ** 2 * D ** 2 * mod
        sel = inr.copy()
        sel[0, 0, 0] = False
        assert sel.sum() > 100
        assert np.allclose(I[sel], expect[sel], rtol=1e-10, atol=1e-12 * expect.max())
        even = (mx[:, None, None] % 2 == 0) & (my[None, :, None] % 2 == 0) & (mz[None, None, :] % 2 == 0)
        assert np.all(np.abs(I[sel & even]) <= 1e-20 * expect.max())


def test_solvent_pads_closed_form():
    """0.998 g/cm^3 water: n_w = 0.998e-24 V / 18.015 * 6.022e23 = 0.03336 V
    molecules (0.3336 e/A^3 at 10 e per molecule); per voxel of a 64^3 cell in
    a 37 A box, O = n_w / 64^3 and H = 2 O; ions count / 64^3."""
    pads = O.solvent_pads((37.0, 37.0, 37.0), (64, 64, 64), "SPCE", {"Na": 10, "Cl": 12})
    n_w = 0.998e-24 * 37.0 ** 3 / 18.015 * 6.022e23
    assert abs(n_w / 37.0 ** 3 - 0.03336) < 1e-5
    assert pads["O"] == pytest.approx(n_w / 64 ** 3, rel=1e-6)
    assert pads["H"] == pytest.approx(2 * n_w / 64 ** 3, rel=1e-6)
    assert pads["Na"] == pytest.approx(10 / 64 ** 3, rel=1e-7)
    assert pads["Cl"] == pytest.approx(12 / 64 ** 3, rel=1e-7)
    with pytest.raises(ValueError):
        O.solvent_pads((37.0,) * 3, (64,) * 3, "NOPE", {})


def test_on_node_atoms_direct_structure_factor_orthorhombic():
    """As above on a 16 x 32 x 64 A box with an 8 x 16 x 32 cell grid (unequal
    reciprocal spacings s_a = 2 pi / (L_a sigma); powers of two keep the
    float32 cell coordinates of the nodes exact): q = sqrt(sum (m_a s_a)^2)
    and the phase sum_a m_a g_a / nn_a per axis."""
    Ls = (16.0, 32.0, 64.0)
    n3 = (8, 16, 32)
    box = np.diag(Ls)
    plan = O.make_plan(n3, box, 4, 2.0, None, 0.0, 100.0)
    assert not plan.isotropic and plan.nn == (16, 32, 64)
    d = plan.border
    rng = np.random.default_rng(6)
    nodes = np.stack([rng.integers(4, n3[0], 5), rng.integers(4, n3[1], 5),
                      rng.integers(d + 5, n3[2] - d, 5)], axis=1)
    pos = (nodes * (np.array(Ls) / np.array(n3))).astype(np.float32)
    syms = ["C", "O", "N", "O", "C"]
    symbols, ids = O.element_ids(syms)
    I, inr, bins = O.frame_intensity(plan, pos, ids, symbols, box)
    mx, my, mz = O.half_spectrum_m(plan.nn)
    s = [float(v) for v in plan.hp.spacing]
    ii, jj, kk = np.nonzero(inr)
    assert len(ii) > 1000
    for t in range(0, len(ii), max(1, len(ii) // 60)):
        i, j, k = ii[t], jj[t], kk[t]
        m = np.array([mx[i], my[j], mz[k]])
        q = math.sqrt(sum((m[a] * s[a]) ** 2 for a in range(3)))
        A = 0j
        for a, sym in zip(nodes, syms):
            A += O.form_factor(sym, q) * np.exp(-2j * np.pi * sum(m[x] * a[x] / plan.nn[x]
                                                                   for x in range(3)))
        lam2 = 1.0
        for ax in range(3):
            x = math.pi * m[ax] / plan.nn[ax]
            if m[ax] != 0:
                g = lambda r: 1 + sum((x / (x + math.pi * jj_)) ** r + (x / (x - math.pi * jj_)) ** r
                                      for jj_ in range(1, 51))
                lam2 *= (g(4) / g(8)) ** 2
        expect = abs(A) ** 2 * lam2
        assert abs(I[i, j, k] - expect) <= 1e-9 * max(1.0, expect), (m, I[i, j, k], expect)


def test_orthorhombic_counts_match_reference_loop():
    """Counts on an orthorhombic grid against a literal enumeration of the
    reference's loop (q^2 = sum (m_a s_a)^2, exact rationals of the float32
    spacings, count 2 per half-spectrum point and 1 at the origin)."""
    Ls, n3 = (30.0, 41.5, 52.0), (32, 32, 32)
    plan = O.make_plan(n3, np.diag(Ls), 4, 2.0, None, 0.0, 0.9)
    S = [Fraction(float(v)) for v in plan.hp.spacing]
    D, Q = Fraction(float(plan.hp.dq)), Fraction(float(plan.hp.qcut))
    C = [0] * plan.hp.num_bins
    M = [int(float(Q / S[a])) + 1 for a in range(3)]
    for i in range(-M[0], M[0] + 1):
        for j in range(-M[1], M[1] + 1):
            for k in range(0, M[2] + 1):
                q2 = i * i * S[0] ** 2 + j * j * S[1] ** 2 + k * k * S[2] ** 2
                if q2 > Q * Q:
                    continue
                b = 0
                while (b + 1) ** 2 * D * D <= q2:
                    b += 1
                if b < plan.hp.num_bins:
                    C[b] += 1 if (i, j, k) == (0, 0, 0) else 2
    pos, syms = O.synthetic_system(200, 30.0, seed=3)
    symbols, ids = O.element_ids(syms)
    H, Cg = O.frame_histogram(plan, pos * np.array(Ls) / 30.0, ids, symbols, np.diag(Ls))
    assert list(Cg.astype(int)) == C
    assert sum(C) > 500


TRIC_BOX = np.array([[16.0, 0.0, 0.0], [8.0, 16.0, 0.0], [4.0, 8.0, 32.0]])


def test_triclinic_inverse_is_exact_inverse():
    """u = M x inverts x = B^T u: M B^T = I (dyadic box: exactly)."""
    M = O.inverse_box_matrix(TRIC_BOX)
    assert np.array_equal(M.astype(np.float64) @ TRIC_BOX.T, np.eye(3))
    rng = np.random.default_rng(1)
    B = np.tril(rng.uniform(5, 50, size=(3, 3)))
    np.fill_diagonal(B, rng.uniform(30, 60, size=3))
    M = O.inverse_box_matrix(B).astype(np.float64)
    assert np.allclose(M @ B.astype(np.float32).T, np.eye(3), atol=1e-6)
    with pytest.raises(ValueError):
        O.inverse_box_matrix(B.T)


def test_on_node_atoms_direct_structure_factor_triclinic():
    """Triclinic cell a = (16,0,0), b = (8,16,0), c = (4,8,32) A on an
    8 x 8 x 16 grid: atoms on grid nodes x = B^T g/n (exact in float32) must
    give |sum_j f_j(|q|) e^{-2 pi i sum_a m_a g_a / nn_a}|^2 prod lambda^2 with
    q = sum_a m_a b_a, b_a = 2 pi (M^T)_a n_a / nn_a, binned by |q| of each
    point itself."""
    n3 = (8, 8, 16)
    plan = O.make_plan(n3, TRIC_BOX, 4, 2.0, None, 0.0, 100.0)
    assert plan.triclinic and plan.nn == (16, 16, 32)
    d = plan.border
    rng = np.random.default_rng(8)
    nodes = np.stack([rng.integers(4, n3[0], 5), rng.integers(4, n3[1], 5),
                      rng.integers(d + 5, n3[2] - d, 5)], axis=1)
    pos = ((nodes / np.array(n3)) @ TRIC_BOX).astype(np.float32)
    syms = ["C", "O", "N", "O", "C"]
    symbols, ids = O.element_ids(syms)
    I, inr, bins = O.frame_intensity(plan, pos, ids, symbols, TRIC_BOX)
    M = O.inverse_box_matrix(TRIC_BOX).astype(np.float64)
    mx, my, mz = O.half_spectrum_m(plan.nn)
    ii, jj, kk = np.nonzero(inr)
    assert len(ii) > 1000
    for t in range(0, len(ii), max(1, len(ii) // 80)):
        i, j, k = ii[t], jj[t], kk[t]
        m = np.array([mx[i], my[j], mz[k]], dtype=np.float64)
        qv = 2 * np.pi * (M.T @ (m * np.array(n3) / np.array(plan.nn)))
        q = float(np.linalg.norm(qv))
        A = 0j
        for g, sym in zip(nodes, syms):
            A += O.form_factor(sym, q) * np.exp(-2j * np.pi * sum(m[a] * g[a] / plan.nn[a]
                                                                   for a in range(3)))
        lam2 = 1.0
        for ax in range(3):
            x = math.pi * m[ax] / plan.nn[ax]
            if m[ax] != 0:
                gsum = lambda r: 1 + sum((x / (x + math.pi * jj_)) ** r + (x / (x - math.pi * jj_)) ** r
                                         for jj_ in range(1, 51))
                lam2 *= (gsum(4) / gsum(8)) ** 2
        expect = abs(A) ** 2 * lam2
        assert abs(I[i, j, k] - expect) <= 1e-6 * max(1.0, expect), (m, I[i, j, k], expect)
        # and the bin is that of |q| itself
        b = bins[i, j, k]
        assert b * float(plan.hp.dq) <= q * (1 + 1e-6) and q < (b + 1) * float(plan.hp.dq) * (1 + 1e-6)


def test_triclinic_counts_match_enumeration():
    """Counts on a triclinic grid: every half-spectrum point with |q| <= qcut
    counts 2 (the origin 1) in the bin of its own |q|, exact rationals."""
    B = np.array([[30.0, 0, 0], [7.5, 28.0, 0], [-5.0, 6.25, 33.0]])
    n3 = (24, 24, 24)
    plan = O.make_plan(n3, B, 4, 2.0, None, 0.0, 0.8)
    bas = [[Fraction(float(v)) for v in row] for row in plan.basis]
    D, Q = Fraction(float(plan.hp.dq)), Fraction(float(plan.hp.qcut))
    C = [0] * plan.hp.num_bins
    mx, my, mz = O.half_spectrum_m(plan.nn)
    R = 12
    for i in mx[np.abs(mx) <= R]:
        for j in my[np.abs(my) <= R]:
            for k in mz[np.abs(mz) <= R]:
                qv = [i * bas[0][c] + j * bas[1][c] + k * bas[2][c] for c in range(3)]
                q2 = sum(v * v for v in qv)
                if q2 > Q * Q:
                    continue
                b = 0
                while (b + 1) ** 2 * D * D <= q2:
                    b += 1
                if b < plan.hp.num_bins:
                    C[b] += 1 if (i, j, k) == (0, 0, 0) else 2
    pos, syms = O.synthetic_system(100, 1.0, seed=4)
    symbols, ids = O.element_ids(syms)
    _, Cg = O.frame_histogram(plan, (pos @ B).astype(np.float32), ids, symbols, B)
    assert list(Cg.astype(int)) == C and sum(C) > 300


# ---- border-slab pad (S2) by closed form -----------------------------------
#
# saxsDeviceKernels.cu:446-465: a voxel is in the border when NOT
# (z > d && z < n_z - d), i.e. z <= d or z >= n_z - d, over every x and y
# (the x / y terms are commented out), and the pad is the voxel sum divided
# by the voxel count n_x n_y (2d + 1).  A lone atom deposits sum_{x,y}
# th_x th_y = 1 over each z plane of its stencil, so the pad is the sum of
# the z weights M_4(w + 3 - o) that land on border planes, over the count.


def _border_expect(nodes_z, w_z, n, d, p=4):
    """Closed-form pad for atoms with cell index nodes_z and fraction w_z."""
    nx, ny, nz = n
    border = set(range(0, d + 1)) | set(range(nz - d, nz))
    total = 0.0
    for mz, w in zip(nodes_z, w_z):
        for o in range(p):
            z = mz - p + o
            if z < 0:
                z += nz
            if z in border:
                total += cardinal_bspline(w + p - 1 - o, p)
    return total / (nx * ny * (2 * d + 1))


@pytest.mark.parametrize("nz", [32, 64, 128])
def test_border_pad_closed_form_straddling_atoms(nz):
    """Atoms whose z stencils straddle z = d / d + 1 and n_z - d - 1 / n_z - d,
    and one that wraps through z = 0: the oracle's spread + border mean must
    equal the closed form.  The wrong slab (z < d, or z <= d without the
    upper slab) or the all-faces border changes the value."""
    n = (16, 16, nz)
    L = (16.0, 16.0, float(nz))                 # 1 A cells: u = x / L exactly
    oc = np.array([1.0 / L[0], 1.0 / L[1], 1.0 / L[2]], dtype=np.float32)
    d = O.border_planes(nz)
    # stencil of cell index m covers m-4 .. m-1; pick m so that it straddles
    # the lower edge (d, d+1), the upper edge (nz-d-1, nz-d) where d >= 2
    # lets a stencil reach it, and wraps through z = 0 (m = 2, 1)
    ms = [d + 3, d + 4, min(nz - 1, nz - d + 1), nz - 1, 2, 1]
    fr = [0.25, 0.625, 0.5, 0.125, 0.375, 0.75]  # exact binary fractions
    xy = [(5.5, 7.25), (9.0, 3.5), (11.75, 12.5), (2.5, 8.0), (6.0, 6.0), (0.5, 15.5)]
    pos = np.array([[x, y, m + f] for (x, y), m, f in zip(xy, ms, fr)], dtype=np.float32)
    m, w = O.grid_coords(pos, oc, n)
    assert list(m[:, 2]) == ms and np.allclose(w[:, 2], fr, atol=0)
    g = O.spread(pos, oc, n, 4)
    pad = O.border_pad(g, d)
    expect = _border_expect(ms, fr, n, d)
    assert expect > 0
    assert abs(pad - expect) <= 1e-15 * max(1.0, expect), (pad, expect)
    # each plausible mistake gives a different value on these atoms
    wrong_lt = np.concatenate([g[:, :, :d], g[:, :, nz - d:]], axis=2).mean()
    faces = np.ones(n, dtype=bool)
    faces[d + 1:n[0] - d, d + 1:n[1] - d, d + 1:nz - d] = False
    wrong_faces = g[faces].mean()
    assert abs(wrong_lt - expect) > 1e-6 * expect
    assert abs(wrong_faces - expect) > 1e-6 * expect


def test_border_pad_full_lattice_is_one_and_spectrum_vanishes():
    """One atom on every node: the density is 1 everywhere (partition of
    unity), so the pad is exactly 1, the literal embedding is a constant
    nn^3 grid and every q != 0 of the spectrum vanishes."""
    n3 = (16, 16, 16)
    L = 32.0
    box = np.diag([L] * 3)
    plan = O.make_plan(n3, box, 4, 2.0, None, 0.0, 100.0)
    g = np.stack(np.meshgrid(*[np.arange(16)] * 3, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = (g * (L / 16)).astype(np.float32)
    oc = np.array([1.0 / L] * 3, dtype=np.float32)
    rho = O.spread(pos, oc, n3, 4)
    assert np.allclose(rho, 1.0, atol=1e-13, rtol=0)
    assert abs(O.border_pad(rho, plan.border) - 1.0) < 1e-13
    syms = ["C"] * len(pos)
    symbols, ids = O.element_ids(syms)
    I, inr, _ = O.frame_intensity(plan, pos, ids, symbols, box, literal=True)
    I0 = I[0, 0, 0]
    off = inr.copy()
    off[0, 0, 0] = False
    assert I0 > 1e10 and I[off].max() <= 1e-20 * I0


# ---- q axis of the reference's own stored profiles ---------------------------
#
# pycusaxs/data/reference_solvents.db holds five profiles written by the
# reference (fixture tests/golden/reference_solvents_db.json): for each, the
# cell grid, the supercell scale sigma, the requested bin size and cutoff,
# the number of rows and the first ten rows.  The rows are b = 1 .. num_bins-1
# with a non-zero count, q = b * dq printed %10.5f (saxsKernel.cu:210-225),
# num_bins = int(qcut / dq) + 1 (:247) after the reset of :489-526.  Rows 4
# and 5 (64^3, sigma 12, 90 A) exercise the qcut reset: 3.7 -> dq (nf - 1).


def _db_rows():
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_solvents_db.json")
    with open(path) as fh:
        return json.load(fh)["rows"]


@pytest.mark.parametrize("row", range(5))
def test_hist_params_reproduce_reference_database_q_axis(row):
    from pycusaxs_amd import geometry as G
    r = _db_rows()[row]
    n = r["grid_size"]
    L = r["box"]
    sigma = r["supercell_scale"]
    nn = [O.find_closest_product(v, sigma) for v in n]
    assert nn == [G.fft_size(v, sigma) for v in n]
    oc = [np.float32(1.0) / np.float32(v) for v in L]
    hp = O.hist_params(n, nn, oc, sigma, r["bin_size"], r["qcut"])
    hg = G.hist_params(n, nn, oc, sigma, r["bin_size"], r["qcut"])
    assert (float(hp.dq), float(hp.qcut), hp.num_bins) == (hg.dq, hg.qcut, hg.num_bins)
    # every bin 1 .. num_bins-1 holds the shell |m| = b on an axis (dq was
    # reset to the grid spacing), so the stored row count is num_bins - 1
    assert float(hp.dq) > r["bin_size"]
    assert hp.num_bins - 1 == r["n_points"], (hp.num_bins, r["n_points"])
    qs = ["%10.5f" % float(np.float32(np.float32(b) * hp.dq)) for b in range(1, 11)]
    assert [float(s) for s in qs] == [pt[0] for pt in r["first_points"]]
    # the oracle's exact bin rule puts a point in each of those bins
    counts = np.zeros(hp.num_bins, dtype=np.int64)
    K = int(float(hp.qcut) / float(hp.spacing[0])) + 1
    for nsq in range(K * K + 1):
        b = O.bin_of_n(nsq, hp.spacing[0], hp.dq, hp.qcut, hp.num_bins)
        if b >= 0:
            counts[b] += 1
    assert np.count_nonzero(counts[1:]) == r["n_points"]


def test_reference_database_qcut_reset_is_exercised():
    """Rows 4 and 5 are the ones whose cutoff the grid lowers: a reset with
    nf instead of nf - 1, or none at all, changes their row count."""
    r = _db_rows()[3]
    n, L, sigma = r["grid_size"], r["box"], r["supercell_scale"]
    nn = [O.find_closest_product(v, sigma) for v in n]
    oc = [np.float32(1.0) / np.float32(v) for v in L]
    hp = O.hist_params(n, nn, oc, sigma, r["bin_size"], r["qcut"])
    assert float(hp.qcut) < r["qcut"]
    dq = hp.dq
    nf = nn[0] // 2
    wrong_nf = int(np.float32(np.float32(dq * np.float32(nf)) / dq)) + 1
    wrong_none = int(np.float32(np.float32(r["qcut"]) / dq)) + 1
    assert wrong_nf - 1 != r["n_points"] and wrong_none - 1 != r["n_points"]
