"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

CPU: the oracle and the synthetic input generator still reproduce the
committed 5,000-atom histograms; the oracle and the engine's host tables
agree with the committed outputs of the reference's own host helpers
(compiled from its sources when the fixtures were made), so these pins hold
where the reference tree is absent.  The GPU counterpart is
tests/test_gpu_golden.py."""
import json
import os

import numpy as np
import pytest

from oracle import saxs_oracle as O
from pycusaxs_amd import geometry as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="module")
def helpers():
    with open(os.path.join(GOLDEN, "ref_helpers.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", ["atoms5k_box37_avg.npz", "atoms5k_box37_given.npz"])
def test_5k_atom_fixture_reproduced_by_oracle(name):
    d = load(name)
    L, n = float(d["box_len"]), int(d["n"])
    pos, syms = O.synthetic_system(5000, L, seed=1)
    symbols, ids = O.element_ids(syms)
    assert np.array_equal(pos.astype(np.float32), d["positions"])
    assert np.array_equal(ids, d["element_ids"].astype(ids.dtype))
    assert list(symbols) == [str(s) for s in d["symbols"]]
    pads = None
    if bool(d["given"]):
        pads = O.solvent_pads((L, L, L), (n, n, n), "SPCE", {})
        assert np.array_equal([pads.get(s, 0.0) for s in symbols], d["pads"])
    box = np.diag([L] * 3)
    plan = O.make_plan((n,) * 3, box, int(d["order"]), float(d["sigma"]), None, float(d["dq"]),
                       float(d["qcut"]), pad_given=pads)
    H, C = O.frame_histogram(plan, d["positions"], d["element_ids"], symbols, box)
    assert np.array_equal(C, d["C"])
    assert list(C[1:].astype(int)) == [34, 82, 178, 258, 402]
    assert np.allclose(H, d["H"], rtol=1e-12, atol=0)
    rows = np.array(O.saxs_rows(plan, H, C))
    assert np.array_equal(rows[:, 0], d["rows"][:, 0])
    assert np.allclose(rows[:, 1], d["rows"][:, 1], rtol=1e-12)


def test_given_pad_moves_only_odd_shells():
    """The two fixtures differ by (rho_border - pad) * FFT(1_cell), which is zero
    wherever one component of m is even and nonzero (nn = 2n): bin 2 holds
    only such shells here (|m|^2 = 4..8) and is identical; bin 1 is not."""
    a, g = load("atoms5k_box37_avg.npz"), load("atoms5k_box37_given.npz")
    assert a["H"][2] == pytest.approx(g["H"][2], rel=1e-12)
    assert abs(a["H"][1] - g["H"][1]) > 0.5 * a["H"][1]


@pytest.mark.parametrize("p", ["4", "6"])
def test_bspline_against_reference_vectors(helpers, p):
    h = helpers["bspline"][p]
    for w, v in zip(h["w"], h["values"]):
        w32 = float(np.float32(w))
        assert np.allclose(v["theta"], O.bspline_theta(w32, int(p)), atol=3e-7)
        assert np.allclose(v["dtheta"], O.bspline_dtheta(w32, int(p)), atol=3e-7)


def test_moduli_against_reference_vectors(helpers):
    for case in helpers["bspmod"]:
        for N, mod in zip(case["sizes"], case["mod"]):
            assert np.allclose(mod, O.bsp_moduli(N, case["order"]), rtol=2e-5)
            assert np.allclose(mod, G.bspline_moduli(N, case["order"]), rtol=2e-5)


def test_cell_inverse_against_reference_vectors(helpers):
    for case in helpers["cell"]:
        oc = np.array(case["oc"]).reshape(3, 3)
        box = np.diag(case["lengths"])
        assert np.array_equal(np.diag(oc).astype(np.float32), G.inverse_box_diag(box))
        assert np.array_equal(np.diag(oc).astype(np.float32), O.inverse_box(box))


def test_solvent_pads_against_reference_vectors(helpers):
    for case in helpers["solvent"]:
        mine = G.explicit_solvent_pads(case["lengths"], case["n"], case["model"], case["ions"])
        orc = O.solvent_pads(case["lengths"], case["n"], case["model"], case["ions"])
        assert set(mine) == set(orc) == set(case["pads"])
        for s, v in case["pads"].items():
            assert np.float32(mine[s]) == np.float32(v)
            assert np.float32(orc[s]) == np.float32(v)


def test_triclinic_inverse_against_reference_vectors(helpers):
    for case in helpers["triclinic"]:
        M_ref = np.array(case["inverse"], dtype=np.float64).reshape(3, 3).astype(np.float32)
        assert np.array_equal(M_ref, G.inverse_box_matrix(case["box"]))
        assert np.array_equal(M_ref, O.inverse_box_matrix(case["box"]))
