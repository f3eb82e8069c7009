"""NVT mode on the CPU: the oracle's restatement (spectrum summed over the
frames, binned once with the first frame's box, times 1/F) pinned by two
closed forms, and the host count table the engine's NVT rows divide by
(geometry.half_spectrum_counts) against the oracle's counts, which
test_oracle_pins.py pins by enumeration; and the engine's NVT q tables,
applied on the CPU to the oracle's per-element spectra as the kernels apply
them, against the oracle's NVT restatement (cubic, orthorhombic and
triclinic boxes changing between frames).

Closed forms: (1) for a fixed box NVT gives the NPT rows, since the binning
is linear and the counts of F equal binnings are F times one binning's;
(2) with the dummy element (f = 1) a frame and a copy of it scaled by
lambda > 1 together with its box have the same grid density, so the same
I(m) at every m whose q <= qcut in the first box (the larger box keeps
every such m): binned with the first box, the two-frame NVT rows are the
first frame's own rows.  Binning with the last box, or dividing by the
wrong frame count, fails (2)."""
import numpy as np
import pytest

from oracle import saxs_oracle as O
from pycusaxs_amd import geometry as G


def _plan_of(n3, plan0, qcut):
    def plan_of(box):
        p = O.make_plan(n3, box, 4, 2.0, None, 0.0, qcut)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        return p
    return plan_of


def test_nvt_fixed_box_equals_npt():
    L, n3, qcut = 30.0, (24, 24, 24), 0.9
    base, syms = O.synthetic_system(400, L, seed=31)
    symbols, ids = O.element_ids(syms)
    box = np.diag([L] * 3)
    frames = [(O.frame_positions(base, 31, f, 0.6), box) for f in range(3)]
    plan = O.make_plan(n3, box, 4, 2.0, None, 0.0, qcut)
    Hn, Cn = O.trajectory_histogram_nvt(plan, frames, ids, symbols, _plan_of(n3, plan, qcut))
    Hp, Cp = O.trajectory_histogram(plan, frames, ids, symbols)
    assert np.array_equal(Cn * 3, Cp)
    ok = Cp > 0
    assert ok.sum() >= 8
    np.testing.assert_allclose((Hn / np.where(ok, Cn, 1))[ok], (Hp / np.where(ok, Cp, 1))[ok],
                               rtol=1e-12)


def test_nvt_bins_with_the_first_box():
    L, lam, n3, qcut = 30.0, 1.125, (24, 24, 24), 0.9
    rng = np.random.default_rng(32)
    pos = rng.uniform(0.0, L, size=(300, 3)).astype(np.float32)
    symbols, ids = ("Du",), np.zeros(300, dtype=np.int32)
    boxA, boxB = np.diag([L] * 3), np.diag([L * lam] * 3)
    posB = (pos.astype(np.float64) * lam).astype(np.float32)
    plan0 = O.make_plan(n3, boxA, 4, 2.0, None, 0.0, qcut)
    plan_of = _plan_of(n3, plan0, qcut)
    Hn, Cn = O.trajectory_histogram_nvt(plan0, [(pos, boxA), (posB, boxB)], ids, symbols, plan_of)
    Ha, Ca = O.frame_histogram(plan0, pos, ids, symbols, boxA)
    assert np.array_equal(Cn, Ca)
    ok = Ca > 0
    np.testing.assert_allclose(Hn[ok], Ha[ok], rtol=2e-6)
    # the other order bins with the larger box: different bins, other rows
    Hr, Cr = O.trajectory_histogram_nvt(plan_of(boxB), [(posB, boxB), (pos, boxA)], ids,
                                        symbols, plan_of)
    assert not np.array_equal(Cr, Ca)


def _tables(n3, box, qcut):
    """The engine's q tables of one box, built as SaxsEngine._q_tables builds
    them (host code only)."""
    nn = tuple(G.fft_size(v, 2.0) for v in n3)
    cell = G.inverse_box_matrix(box)
    h = G.hist_params(n3, nn, np.diag(cell).astype(np.float32), 2.0, 0.0, qcut)
    s = h.spacing
    if G.is_triclinic(cell):
        basis = G.reciprocal_basis(cell, n3, nn)
        kmax = G.triclinic_kmax(basis, h.qcut, nn)
        return nn, h, G.triclinic_bins(basis, h.dq, h.qcut, h.num_bins, kmax), kmax, True
    if s[0] == s[1] == s[2]:
        return nn, h, G.shell_bins(s[0], h.dq, h.qcut, h.num_bins), None, False
    kmax = tuple(min(k, v // 2) for k, v in zip(G.point_kmax(s, h.qcut), nn))
    return nn, h, G.point_bins(s, h.dq, h.qcut, h.num_bins, kmax), kmax, False


@pytest.mark.parametrize("box", [np.diag([30.0] * 3), np.diag([30.0, 41.5, 52.0]),
                                 np.array([[30.0, 0, 0], [7.5, 28.0, 0], [-5.0, 6.25, 33.0]])],
                         ids=["cubic", "orthorhombic", "triclinic"])
def test_half_spectrum_counts_match_oracle(box):
    n3, qcut = (24, 24, 24), 0.8
    nn, h, bins, kmax, signed = _tables(n3, box, qcut)
    got = G.half_spectrum_counts(nn, bins, kmax, h.num_bins, signed=signed)
    plan = O.make_plan(n3, box, 4, 2.0, None, 0.0, qcut)
    assert plan.hp.num_bins == h.num_bins
    bo, inr, _ = O.spectrum_bins(plan)
    _, Co = O.histogram(plan, np.zeros(bo.shape), inr, bo)
    assert np.array_equal(got, Co) and Co.sum() > 300


NVT_BOXES = {
    "cubic": [np.diag([30.0] * 3), np.diag([30.5] * 3), np.diag([29.6] * 3)],
    "orthorhombic": [np.diag([30.0] * 3), np.diag([30.4, 29.8, 30.2]), np.diag([29.7, 30.0, 30.3])],
    "triclinic": [np.diag([30.0] * 3), np.array([[30.0, 0, 0], [2.5, 29.5, 0], [-1.5, 1.0, 30.4]]),
                  np.array([[30.3, 0, 0], [2.0, 29.8, 0], [-1.0, 1.5, 30.1]])],
}


@pytest.mark.parametrize("kind", sorted(NVT_BOXES))
def test_binned_tables_reproduce_oracle_nvt(kind):
    """The engine's NVT tables (SaxsEngine._q_tables_binned: frame 0's bins,
    the frame's own form factors and cut) and count table, applied on the
    CPU to the oracle's per-element spectra the way the kernels apply them
    (A = sum_e f_e[key] F_e at keyed points, 2 I per point, I at the origin,
    into the bin of the key), against the oracle's NVT restatement."""
    import types
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec
    boxes = [np.asarray(b, dtype=np.float32) for b in NVT_BOXES[kind]]
    n3, qcut = (24, 24, 24), 0.9
    base, syms = O.synthetic_system(400, 30.0, seed=33)
    frames = [O.frame_positions(base, 33, f, 0.5) for f in range(len(boxes))]
    symbols, ids = O.element_ids(syms)
    plan0 = O.make_plan(n3, boxes[0], 4, 2.0, None, 0.0, qcut)
    plan_of = _plan_of(n3, plan0, qcut)
    H, C = O.trajectory_histogram_nvt(plan0, list(zip(frames, boxes)), ids, symbols, plan_of)

    nn = plan0.nn
    fe = types.SimpleNamespace(spec=SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, qcut))
    cell0 = G.inverse_box_matrix(boxes[0])
    fe.hist = G.hist_params(n3, nn, np.diag(cell0).astype(np.float32), 2.0, 0.0, qcut)
    fe._q_tables = lambda h, c: SaxsEngine._q_tables(fe, h, c)
    fe._hist_for = lambda c: SaxsEngine._hist_for(fe, c)
    He = np.zeros(fe.hist.num_bins)
    X, Y, Z = np.meshgrid(*G.half_spectrum_m(nn), indexing="ij")
    w = np.where((X == 0) & (Y == 0) & (Z == 0), 1.0, 2.0)
    for pos, box in zip(frames, boxes):
        cell = G.inverse_box_matrix(box)
        h = fe._hist_for(cell)
        if np.array_equal(cell, cell0):
            bins, ff, kmax = fe._q_tables(h, cell)
        else:
            bins, ff, kmax = SaxsEngine._q_tables_binned(fe, h, cell, fe._hist_for(cell0), cell0)
        K = (int(np.sqrt(len(bins) - 1)),) * 3 if kmax is None else kmax
        valid = (np.abs(X) <= K[0]) & (np.abs(Y) <= K[1]) & (np.abs(Z) <= K[2])
        if kmax is None:
            key = X * X + Y * Y + Z * Z
            valid &= key < len(bins)
        elif G.is_triclinic(cell):
            key = ((X + K[0]) * (2 * K[1] + 1) + Y + K[1]) * (2 * K[2] + 1) + Z + K[2]
        else:
            key = (np.abs(X) * (K[1] + 1) + np.abs(Y)) * (K[2] + 1) + np.abs(Z)
        key = np.where(valid, key, 0)
        p = plan_of(box)
        Fs, _, _, _ = O.frame_spectrum_terms(p, pos, ids, symbols, box)
        A = sum(np.where(valid, ff[e][key], 0.0) * Fs[e] for e in range(len(symbols)))
        I = np.abs(A) ** 2 * p.moduli[0][:, None, None] * p.moduli[1][None, :, None] \
            * p.moduli[2][None, None, :A.shape[2]]
        b = np.where(valid, bins[key], -1)
        np.add.at(He, b[b >= 0], (w * I)[b >= 0])
    bins0, _, kmax0 = fe._q_tables(fe.hist, cell0)
    Ce = G.half_spectrum_counts(nn, bins0, kmax0, fe.hist.num_bins, signed=G.is_triclinic(cell0))
    assert np.array_equal(Ce, C)
    ok = C > 0
    np.testing.assert_allclose(He[ok] / len(boxes), H[ok], rtol=1e-6)
