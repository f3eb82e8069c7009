"""Pins the oracle's analytic adjoint (grad_linear_H) against central finite
differences of the oracle's own forward pass, in float64 throughout
(grid coordinates included), on small systems."""
import numpy as np
import pytest

from oracle import saxs_oracle as O


@pytest.fixture
def exact_coords():
    O.EXACT_COORDS = True
    yield
    O.EXACT_COORDS = False


@pytest.mark.parametrize("order,qcut", [(4, 100.0), (6, 100.0), (4, 0.9)])
def test_adjoint_matches_central_differences(exact_coords, order, qcut):
    L, n = 20.0, 16
    box = np.diag([L] * 3)
    rng = np.random.default_rng(order)
    pos = rng.uniform(0, L, size=(40, 3))
    syms = list(rng.choice(["C", "O", "N", "H"], size=40))
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((n,) * 3, box, order, 2.0, None, 0.0, qcut)
    H0, C0 = O.frame_histogram(plan, pos, ids, symbols, box, literal=False)
    gH = rng.normal(size=H0.shape) / np.maximum(np.abs(H0), 1.0)
    grad = O.grad_linear_H(plan, pos, ids, symbols, box, gH)

    def loss(p):
        H, _ = O.frame_histogram(plan, p, ids, symbols, box, literal=False)
        return float(gH @ H)

    h = 1e-5
    for i in rng.choice(len(pos), size=6, replace=False):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert abs(fd - grad[i, a]) <= 1e-6 * max(1.0, np.abs(grad).max()), (i, a, fd, grad[i, a])


def test_pad_term_matters(exact_coords):
    """Atoms near the border slabs feel the bulk-solvent mean: dropping the
    pad-mean term from the adjoint must break the finite-difference match."""
    L, n = 20.0, 16
    box = np.diag([L] * 3)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, L, size=(30, 3))
    pos[:10, 2] = rng.uniform(0, 1.5, size=10)            # close to z = 0 (border)
    syms = ["C"] * 30
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, 100.0)
    H0, _ = O.frame_histogram(plan, pos, ids, symbols, box, literal=False)
    gH = rng.normal(size=H0.shape) / np.maximum(np.abs(H0), 1.0)
    grad = O.grad_linear_H(plan, pos, ids, symbols, box, gH)
    h = 1e-5
    i, a = 0, 2
    pp, pm = pos.copy(), pos.copy()
    pp[i, a] += h
    pm[i, a] -= h
    lp = gH @ O.frame_histogram(plan, pp, ids, symbols, box, literal=False)[0]
    lm = gH @ O.frame_histogram(plan, pm, ids, symbols, box, literal=False)[0]
    fd = (lp - lm) / (2 * h)
    assert abs(fd - grad[i, a]) <= 1e-6 * max(1.0, abs(fd))
    # the border slab really carries density here, so the pad mean is nonzero
    g = O.spread(pos, O.inverse_box(box), plan.n, 4)
    assert O.border_pad(g, plan.border) > 0


def test_given_pad_adjoint_matches_central_differences(exact_coords):
    """padding::given: the pad is a constant, so the adjoint has no pad-mean
    term; central differences of the given-mode forward pin that, with atoms
    placed in the border slabs where the avg-mode term would be nonzero."""
    L, n = 20.0, 16
    box = np.diag([L] * 3)
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, L, size=(30, 3))
    pos[:10, 2] = rng.uniform(0, 1.5, size=10)
    syms = list(rng.choice(["O", "H", "C"], size=30))
    symbols, ids = O.element_ids(syms)
    pads = O.solvent_pads((L, L, L), (n, n, n), "SPCE", {})
    plan = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, 100.0, pad_given=pads)
    H0, _ = O.frame_histogram(plan, pos, ids, symbols, box, literal=False)
    gH = rng.normal(size=H0.shape) / np.maximum(np.abs(H0), 1.0)
    grad = O.grad_linear_H(plan, pos, ids, symbols, box, gH)
    h = 1e-5
    for i in list(range(3)) + list(rng.choice(np.arange(10, 30), size=3, replace=False)):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            lp = gH @ O.frame_histogram(plan, pp, ids, symbols, box, literal=False)[0]
            lm = gH @ O.frame_histogram(plan, pm, ids, symbols, box, literal=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grad[i, a]) <= 1e-6 * max(1.0, np.abs(grad).max()), (i, a, fd, grad[i, a])


def test_triclinic_adjoint_matches_central_differences(exact_coords):
    """The chain rule through u = M x (dL/dx = sum_a dL/dr_a n_a M_a.) and the
    own-bin histogram of triclinic cells, against central differences."""
    B = np.array([[20.0, 0, 0], [5.0, 18.0, 0], [-3.0, 4.0, 22.0]])
    n = 16
    rng = np.random.default_rng(12)
    pos = rng.uniform(0, 1, size=(30, 3)) @ B
    syms = list(rng.choice(["C", "O", "N"], size=30))
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((n,) * 3, B, 4, 2.0, None, 0.0, 100.0)
    assert plan.triclinic
    H0, _ = O.frame_histogram(plan, pos, ids, symbols, B, literal=False)
    gH = rng.normal(size=H0.shape) / np.maximum(np.abs(H0), 1.0)
    grad = O.grad_linear_H(plan, pos, ids, symbols, B, gH)
    h = 1e-5
    for i in rng.choice(len(pos), size=5, replace=False):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            lp = gH @ O.frame_histogram(plan, pp, ids, symbols, B, literal=False)[0]
            lm = gH @ O.frame_histogram(plan, pm, ids, symbols, B, literal=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grad[i, a]) <= 1e-6 * max(1.0, np.abs(grad).max()), (i, a, fd, grad[i, a])
