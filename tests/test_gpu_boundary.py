"""End-to-end: pycusaxs_cuda.run on the MI355X engine with an in-memory
topology writes the same saxs.dat rows as the oracle (q column exact, I(q)
within 1e-5 relative), honours begin/end/stride, and returns the reference's
result dict."""
import numpy as np
import pytest

import pycusaxs_cuda
from oracle import saxs_oracle as O
from tests.fake_topology import FakeTopology

pytestmark = pytest.mark.gpu


def parse(path):
    rows = []
    for line in open(path):
        rows.append((float(line[:10]), float(line[10:22])))
    return rows


def test_run_writes_oracle_rows(gpu, tmp_path):
    L = 37.0
    base, syms = O.synthetic_system(3000, L, seed=21)
    frames = [O.frame_positions(base, 21, f, 0.5) for f in range(7)]
    topo = FakeTopology(frames, syms, np.diag([L] * 3))
    out = tmp_path / "saxs.dat"
    res = pycusaxs_cuda.run(obj_topology=topo, topology="sys.tpr", trajectory="traj.xtc",
                            grid=[48], begin=1, end=6, stride=2, output=str(out), order=4,
                            scale_factor=2.0, bin_size=0.02, qcut=0.8)
    assert res["grid"] == [48, 48, 48] and res["scaled_grid"] == [0, 0, 0]
    assert "Total Frames  : 3" in res["summary"]
    assert topo.requested == [(1, 7, 2)]
    used = [frames[i] for i in (1, 3, 5)]
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((48,) * 3, np.diag([L] * 3), 4, 2.0, None, 0.02, 0.8)
    H, C = O.trajectory_histogram(plan, [(f, np.diag([L] * 3)) for f in used], ids, symbols)
    want = O.saxs_rows(plan, H, C)
    got = parse(out)
    assert len(got) == len(want)
    for (qg, ig), (qw, iw) in zip(got, want):
        assert ("%10.5f" % qg) == ("%10.5f" % qw)
        assert abs(ig - iw) <= 1e-5 * abs(iw) + 1e-5 * 10 ** np.floor(np.log10(abs(iw)))


def test_run_with_scaled_grid_and_nvt(gpu, tmp_path):
    L = 30.0
    pos, syms = O.synthetic_system(1500, L, seed=22)
    topo = FakeTopology([pos], syms, np.diag([L] * 3))
    out = tmp_path / "o.dat"
    res = pycusaxs_cuda.run(topo, "a", "b", [32], scaled_grid=[64], output=str(out),
                            bin_size=0.0, qcut=1.0, simulation="nvt")
    assert res["scaled_grid"] == [64, 64, 64]
    symbols, ids = O.element_ids(syms)
    # Q2: with an explicit scaled grid the reference still resets dq with
    # sigma = scale_factor (default 1.0), not nn/n
    plan = O.make_plan((32,) * 3, np.diag([L] * 3), 4, 1.0, (64, 64, 64), 0.0, 1.0)
    H, C = O.frame_histogram(plan, pos, ids, symbols, np.diag([L] * 3))
    want = O.saxs_rows(plan, H, C)
    got = parse(out)
    assert len(got) == len(want) and len(got) > 0
    for (qg, ig), (qw, iw) in zip(got, want):
        assert abs(ig - iw) <= 2e-5 * abs(iw)


TRIC_A = np.array([[37.0, 0.0, 0.0], [3.0, 36.5, 0.0], [-2.0, 1.5, 37.5]], dtype=np.float32)
TRIC_B = np.array([[37.3, 0.0, 0.0], [2.5, 36.8, 0.0], [-1.5, 2.0, 37.2]], dtype=np.float32)
NVT_BOXES = {
    "cubic": [np.diag([37.0] * 3), np.diag([37.6] * 3), np.diag([36.5] * 3), np.diag([37.2] * 3)],
    "orthorhombic": [np.diag([37.0] * 3), np.diag([37.5, 36.8, 37.3]), np.diag([37.0] * 3),
                     np.diag([36.9, 37.4, 37.0])],
    "triclinic": [np.diag([37.0] * 3), TRIC_A, np.diag([37.3] * 3), TRIC_B],
}


@pytest.mark.parametrize("kind", sorted(NVT_BOXES))
def test_run_nvt_changing_box(gpu, tmp_path, kind):
    """NVT with a box that changes between frames: every frame's spectrum
    (its own q cut and form factors) is summed and binned once with the
    first frame's box, times 1/F, against the oracle's restatement."""
    boxes = [np.asarray(b, dtype=np.float32) for b in NVT_BOXES[kind]]
    base, syms = O.synthetic_system(3000, 37.0, seed=23)
    frames = [O.frame_positions(base, 23, f, 0.5) for f in range(4)]
    topo = FakeTopology(frames, syms, boxes[0])
    topo.boxes = boxes
    out = tmp_path / "nvt.dat"
    pycusaxs_cuda.run(topo, "sys.tpr", "traj.xtc", [48], end=3, output=str(out),
                      scale_factor=2.0, bin_size=0.02, qcut=0.8, simulation="nvt")
    symbols, ids = O.element_ids(syms)
    plan0 = O.make_plan((48,) * 3, boxes[0], 4, 2.0, None, 0.02, 0.8)

    def plan_of(box):
        p = O.make_plan((48,) * 3, box, 4, 2.0, None, 0.02, 0.8)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        return p

    H, C = O.trajectory_histogram_nvt(plan0, list(zip(frames, boxes)), ids, symbols, plan_of)
    want = O.saxs_rows(plan0, H, C)
    got = parse(out)
    assert len(got) == len(want) and len(got) >= 8
    for (qg, ig), (qw, iw) in zip(got, want):
        assert ("%10.5f" % qg) == ("%10.5f" % qw)
        assert abs(ig - iw) <= 1e-5 * abs(iw) + 1e-5 * 10 ** np.floor(np.log10(abs(iw)))


def test_run_rejects_unknown_ensemble(tmp_path):
    topo = FakeTopology([np.zeros((3, 3), np.float32)], ["O", "H", "H"], np.diag([30.0] * 3))
    with pytest.raises(ValueError, match="simulation"):
        pycusaxs_cuda.run(topo, "a", "b", [16], output=str(tmp_path / "x.dat"),
                          scale_factor=2.0, simulation="NVT")


def test_run_given_padding(gpu, tmp_path):
    """padding="given" (keyword-only extension): explicit SPCE water plus the
    sodium/chlorine counts as the bulk-solvent pad, against the oracle with
    the AtomCounter pads of the first frame's box."""
    L = 37.0
    base, syms = O.synthetic_system(3000, L, seed=24)
    frames = [O.frame_positions(base, 24, f, 0.5) for f in range(3)]
    topo = FakeTopology(frames, syms, np.diag([L] * 3))
    out = tmp_path / "g.dat"
    res = pycusaxs_cuda.run(topo, "sys.tpr", "traj.xtc", [48], end=2, output=str(out),
                            scale_factor=2.0, bin_size=0.02, qcut=0.8, water_model="SPCE",
                            sodium=4, chlorine=4, padding="given")
    assert "Padding Mode  : given" in res["summary"]
    symbols, ids = O.element_ids(syms)
    pads = O.solvent_pads((L, L, L), (48, 48, 48), "SPCE", {"Na": 4, "Cl": 4})
    plan = O.make_plan((48,) * 3, np.diag([L] * 3), 4, 2.0, None, 0.02, 0.8, pad_given=pads)
    H, C = O.trajectory_histogram(plan, [(f, np.diag([L] * 3)) for f in frames], ids, symbols)
    want = O.saxs_rows(plan, H, C)
    got = parse(out)
    assert len(got) == len(want) and len(got) > 0
    for (qg, ig), (qw, iw) in zip(got, want):
        assert ("%10.5f" % qg) == ("%10.5f" % qw)
        assert abs(ig - iw) <= 1e-5 * abs(iw) + 1e-5 * 10 ** np.floor(np.log10(abs(iw)))


def test_run_npt_box_growing_past_the_workspace(gpu, tmp_path):
    """NPT with a box that grows 10 % over four frames at qcut 1.6: the kept
    range goes from K = 18 to 20, past the one extra frequency a plan's
    workspace holds.  Every frame is processed (the engine makes a plan for
    the larger box, keeping H and C) and the rows equal the oracle's, each
    frame with its own box's q and the first frame's dq, qcut and bins."""
    L0 = 37.0
    Ls = [37.0, 38.2, 39.5, 40.7]
    base, syms = O.synthetic_system(3000, L0, seed=31)
    frames = [(O.frame_positions(base, 31, f, 0.3) * (L / L0)).astype(np.float32)
              for f, L in enumerate(Ls)]
    boxes = [np.diag([L] * 3).astype(np.float32) for L in Ls]
    topo = FakeTopology(frames, syms, boxes[0])
    topo.boxes = boxes
    out = tmp_path / "npt.dat"
    pycusaxs_cuda.run(topo, "sys.tpr", "traj.xtc", [48], end=3, output=str(out),
                      scale_factor=2.0, bin_size=0.02, qcut=1.6)
    symbols, ids = O.element_ids(syms)
    plan0 = O.make_plan((48,) * 3, boxes[0], 4, 2.0, None, 0.02, 1.6)

    def plan_of(box):
        p = O.make_plan((48,) * 3, box, 4, 2.0, None, 0.02, 1.6)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        return p

    ks = [int(float(plan0.hp.qcut) / float(plan_of(b).hp.spacing[0])) for b in boxes]
    assert ks[0] == 18 and ks[-1] >= 20
    H = np.zeros(plan0.hp.num_bins)
    C = np.zeros(plan0.hp.num_bins)
    for f, b in zip(frames, boxes):
        h, c = O.frame_histogram(plan_of(b), f, ids, symbols, b)
        H += h
        C += c
    want = O.saxs_rows(plan0, H, C)
    got = parse(out)
    assert len(got) == len(want) and len(got) >= 15
    for (qg, ig), (qw, iw) in zip(got, want):
        assert ("%10.5f" % qg) == ("%10.5f" % qw)
        assert abs(ig - iw) <= 1e-5 * abs(iw) + 1e-5 * 10 ** np.floor(np.log10(abs(iw)))


def test_engine_box_growth_makes_a_plan_and_keeps_the_accumulators(gpu):
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    base, syms = O.synthetic_system(2000, 37.0, seed=32)
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((48,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.02, 1.6), np.diag([37.0] * 3),
                     max_atoms=2000, max_frames=1)
    k0 = eng.kept()
    eng.forward(torch.from_numpy(base[None]), torch.from_numpy(ids))
    H0 = eng.H.clone()
    big = (base * (41.0 / 37.0)).astype(np.float32)
    eng.forward(torch.from_numpy(big[None]), torch.from_numpy(ids), box=np.diag([41.0] * 3))
    torch.cuda.synchronize()
    assert eng.plans_created == 2 and eng.kept()[2] > k0[2] + 1
    plan0 = O.make_plan(n3, np.diag([37.0] * 3), 4, 2.0, None, 0.02, 1.6)
    p1 = O.make_plan(n3, np.diag([41.0] * 3), 4, 2.0, None, 0.02, 1.6)
    p1.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p1.hp.spacing)
    h1, c1 = O.frame_histogram(p1, big, ids, symbols, np.diag([41.0] * 3))
    h0, c0 = O.frame_histogram(plan0, base, ids, symbols, np.diag([37.0] * 3))
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    assert np.array_equal(C, c0 + c1)
    sel = (np.arange(len(C)) >= 1) & (C > 0)
    rel = np.abs(H[sel] - (h0 + h1)[sel]) / np.abs((h0 + h1)[sel])
    assert rel.max() <= 1e-5
    assert torch.all(H0[1:] <= torch.from_numpy(H[1:]).to(H0.device) + 1e-9 * H0[1:].abs())


def test_run_nvt_rectangular_after_triclinic(gpu, tmp_path):
    """NVT bins every frame with the first frame's (triclinic) q; a
    rectangular frame after it runs in the same batch with signed point keys
    (per-frame tables on the device), against the oracle's NVT restatement."""
    boxes = [TRIC_A, np.diag([37.0] * 3).astype(np.float32), TRIC_B]
    base, syms = O.synthetic_system(1500, 37.0, seed=33)
    frames = [O.frame_positions(base, 33, f, 0.3) for f in range(3)]
    topo = FakeTopology(frames, syms, boxes[0])
    topo.boxes = boxes
    out = tmp_path / "x.dat"
    pycusaxs_cuda.run(topo, "sys.tpr", "traj.xtc", [48], end=2, output=str(out),
                      scale_factor=2.0, bin_size=0.02, qcut=0.8, simulation="nvt")
    symbols, ids = O.element_ids(syms)
    plan0 = O.make_plan((48,) * 3, boxes[0], 4, 2.0, None, 0.02, 0.8)

    def plan_of(box):
        p = O.make_plan((48,) * 3, box, 4, 2.0, None, 0.02, 0.8)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        return p

    H, C = O.trajectory_histogram_nvt(plan0, list(zip(frames, boxes)), ids, symbols, plan_of)
    want = O.saxs_rows(plan0, H, C)
    got = parse(out)
    assert len(got) == len(want) and len(got) >= 8
    for (qg, ig), (qw, iw) in zip(got, want):
        assert ("%10.5f" % qg) == ("%10.5f" % qw)
        assert abs(ig - iw) <= 1e-5 * abs(iw) + 1e-5 * 10 ** np.floor(np.log10(abs(iw)))


def test_given_pad_above_bound_rejected(gpu):
    """The C entry point rejects a given pad just above 2^31 / n_z atoms per
    voxel (the engine's Python check is bypassed here)."""
    import ctypes
    import numpy as np_
    from pycusaxs_amd import _lib
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    from pycusaxs_amd.geometry import max_given_pad
    n3, nn = resolve_grid((32,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, ("C", "O"), 2.0, 0.0, 1.0), np.diag([30.0] * 3),
                     max_atoms=10, max_frames=1)
    lim = max_given_pad(32)
    ok = np_.array([lim, 0.5], dtype=np_.float64)
    assert eng.lib.saxs_plan_set_padding(eng._plan, ok.ctypes.data_as(ctypes.c_void_p)) == 0
    bad = np_.array([np_.nextafter(lim, np_.inf), 0.5], dtype=np_.float64)
    rc = eng.lib.saxs_plan_set_padding(eng._plan, bad.ctypes.data_as(ctypes.c_void_p))
    assert rc == _lib.SAXS_EINVAL
    with pytest.raises(ValueError):
        eng.set_padding({"C": float(bad[0])})
    eng.close()
