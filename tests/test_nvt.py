"""NVT mode on the CPU: the oracle's restatement (spectrum summed over the
frames, binned once with the first frame's box, times 1/F) pinned by two
closed forms, and the host count table the engine's NVT rows divide by
(geometry.half_spectrum_counts) against the oracle's counts, which
test_oracle_pins.py pins by enumeration.

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
