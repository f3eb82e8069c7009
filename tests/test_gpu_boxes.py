"""Frames whose boxes differ in one launch chain (saxs_forward_boxes): per-frame
cells and per-frame q tables built on the device, against the engine's
per-box path (host tables, one frame per launch: counts equal, H within
float32 rounding) and against the FP64 oracle (counts exact, I(q) within
1e-5), for cubic, orthorhombic, triclinic and mixed boxes, NPT and NVT."""
import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

L0, N_CELL, QCUT, DQ = 37.0, 48, 0.8, 0.02
TA = np.array([[37.0, 0.0, 0.0], [3.0, 36.5, 0.0], [-2.0, 1.5, 37.5]], dtype=np.float32)
TB = np.array([[37.3, 0.0, 0.0], [2.5, 36.8, 0.0], [-1.5, 2.0, 37.2]], dtype=np.float32)
BOXES = {
    "cubic": [np.diag([L]) * np.eye(3) for L in (37.0, 37.4, 36.6, 37.9, 37.0, 36.2)],
    "orthorhombic": [np.diag(v) for v in ((37.0, 37.0, 37.0), (37.5, 36.8, 37.3),
                                          (36.9, 37.4, 37.0), (38.0, 36.5, 37.1),
                                          (37.2, 37.2, 37.2), (36.6, 37.7, 37.4))],
    "triclinic": [TA, TB, TA * np.float32(1.01), TB * np.float32(0.99), TA, TB],
    "mixed": [np.diag([37.0] * 3), TA, np.diag((37.5, 36.8, 37.3)), TB, np.diag([37.6] * 3),
              TA * np.float32(1.02)],
}


def _frames(kind, seed):
    boxes = [np.asarray(b, dtype=np.float32) for b in BOXES[kind]]
    base, syms = O.synthetic_system(2500, 1.0, seed=seed)       # fractional coordinates
    frames = []
    for f, b in enumerate(boxes):
        u = (base + np.random.default_rng([seed, f]).normal(0, 0.004, base.shape)) % 1.0
        frames.append((u.astype(np.float64) @ b.astype(np.float64)).astype(np.float32))
    return boxes, frames, syms


def _engine(symbols, box0, F):
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n3, nn = resolve_grid((N_CELL,) * 3, 2.0)
    return SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, DQ, QCUT), box0,
                      max_atoms=2500, max_frames=F)


def _plan_of(plan0):
    def make(box):
        p = O.make_plan((N_CELL,) * 3, box, 4, 2.0, None, DQ, QCUT)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        return p
    return make


@pytest.mark.parametrize("kind", sorted(BOXES))
def test_npt_batch_of_boxes(gpu, kind):
    import torch
    boxes, frames, syms = _frames(kind, 50 + len(kind))
    symbols, ids = O.element_ids(syms)
    xyz = torch.from_numpy(np.stack(frames))
    idt = torch.from_numpy(ids)
    a = _engine(symbols, boxes[0], len(frames))
    a.forward_boxes(xyz, idt, np.stack(boxes))
    b = _engine(symbols, boxes[0], 1)
    for f, box in enumerate(boxes):
        b.forward(xyz[f:f + 1], idt, box=box)
    torch.cuda.synchronize()
    Ha, Ca = a.H.cpu().numpy(), a.C.cpu().numpy()
    Hb, Cb = b.H.cpu().numpy(), b.C.cpu().numpy()
    assert a.frames_done == len(frames)
    a.close()
    b.close()
    assert np.array_equal(Ca, Cb), (Ca, Cb)
    sel = (np.arange(len(Ca)) >= 1) & (Cb > 0)
    rel = np.abs(Ha[sel] - Hb[sel]) / np.abs(Hb[sel])
    print(f"[{kind}] batch vs per-box path: max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-6
    plan0 = O.make_plan((N_CELL,) * 3, boxes[0], 4, 2.0, None, DQ, QCUT)
    make = _plan_of(plan0)
    Ho = np.zeros(plan0.hp.num_bins)
    Co = np.zeros(plan0.hp.num_bins)
    for f, box in zip(frames, boxes):
        h, c = O.frame_histogram(make(box), f, ids, symbols, box)
        Ho += h
        Co += c
    assert np.array_equal(Ca, Co)
    rel = np.abs(Ha[sel] / Ca[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    print(f"[{kind}] batch vs oracle: max rel {rel.max():.2e} over {sel.sum()} bins", flush=True)
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("kind", ["cubic", "triclinic"])
def test_npt_batch_most_outputs_kept(gpu, kind):
    """A 160-point FFT grid (80^3 cells) at q <= 3.5 keeps more than an
    eighth of the y and x axes, where the y and x passes form every stage-1
    output by radix-16 DFTs (160 = 10 x 16): boxes changing every frame
    (per-frame device tables) against the per-box path and the oracle."""
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n_cell, qcut = 80, 3.5
    boxes, frames, syms = _frames(kind, 90 + len(kind))
    boxes, frames = boxes[:3], frames[:3]
    symbols, ids = O.element_ids(syms)
    xyz = torch.from_numpy(np.stack(frames))
    idt = torch.from_numpy(ids)
    n3, nn = resolve_grid((n_cell,) * 3, 2.0)
    assert nn == (160, 160, 160)

    def engine(F):
        return SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, DQ, qcut), boxes[0],
                          max_atoms=2500, max_frames=F)
    a = engine(len(frames))
    a.forward_boxes(xyz, idt, np.stack(boxes))
    b = engine(1)
    for f, box in enumerate(boxes):
        b.forward(xyz[f:f + 1], idt, box=box)
    torch.cuda.synchronize()
    kx, ky, _ = b.kept()
    Ha, Ca = a.H.cpu().numpy(), a.C.cpu().numpy()
    Hb, Cb = b.H.cpu().numpy(), b.C.cpu().numpy()
    a.close()
    b.close()
    assert 8 * ((kx + 1) // 2) >= 160 and 8 * ((ky + 1) // 2) >= 160, (kx, ky)
    assert np.array_equal(Ca, Cb)
    sel = (np.arange(len(Ca)) >= 1) & (Cb > 0)
    rel = np.abs(Ha[sel] - Hb[sel]) / np.abs(Hb[sel])
    print(f"[{kind}, 160 points] batch vs per-box path: max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-6
    plan0 = O.make_plan((n_cell,) * 3, boxes[0], 4, 2.0, None, DQ, qcut)

    def make(box):
        p = O.make_plan((n_cell,) * 3, box, 4, 2.0, None, DQ, qcut)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        return p
    Ho = np.zeros(plan0.hp.num_bins)
    Co = np.zeros(plan0.hp.num_bins)
    for f, box in zip(frames, boxes):
        h, c = O.frame_histogram(make(box), f, ids, symbols, box)
        Ho += h
        Co += c
    assert np.array_equal(Ca, Co)
    rel = np.abs(Ha[sel] / Ca[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    print(f"[{kind}, 160 points] batch vs oracle: max rel {rel.max():.2e} over {sel.sum()} bins",
          flush=True)
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("kind", ["cubic", "orthorhombic", "mixed"])
def test_nvt_batch_of_boxes(gpu, kind):
    """NVT: every frame binned with frame 0's q (its own cut and form
    factors), against the per-box path and the oracle's NVT restatement."""
    import torch
    boxes, frames, syms = _frames(kind, 70 + len(kind))
    symbols, ids = O.element_ids(syms)
    xyz = torch.from_numpy(np.stack(frames))
    idt = torch.from_numpy(ids)
    a = _engine(symbols, boxes[0], len(frames))
    a.forward_boxes(xyz, idt, np.stack(boxes), bin_box=boxes[0])
    torch.cuda.synchronize()
    Ha = a.H.cpu().numpy()
    rows_a = a.nvt_rows(boxes[0], a.frames_done)
    a.close()
    plan0 = O.make_plan((N_CELL,) * 3, boxes[0], 4, 2.0, None, DQ, QCUT)
    Ho, Co = O.trajectory_histogram_nvt(plan0, list(zip(frames, boxes)), ids, symbols,
                                        _plan_of(plan0))
    want = O.saxs_rows(plan0, Ho, Co)
    assert len(rows_a) == len(want) and len(want) >= 8
    for (qa, ia), (qw, iw) in zip(rows_a, want):
        assert ("%10.5f" % qa) == ("%10.5f" % qw)
        assert abs(ia - iw) <= 1e-5 * abs(iw), (qa, ia, iw)
    if kind != "mixed":
        b = _engine(symbols, boxes[0], 1)
        for f, box in enumerate(boxes):
            b.forward(xyz[f:f + 1], idt, box=box, bin_box=boxes[0])
        torch.cuda.synchronize()
        Hb = b.H.cpu().numpy()
        b.close()
        sel = np.abs(Hb) > 0
        sel[0] = False
        assert np.max(np.abs(Ha[sel] - Hb[sel]) / np.abs(Hb[sel])) <= 1e-6


def test_batch_growing_past_the_workspace(gpu):
    """A batch whose boxes grow 10 % (kept range 18 -> 20 at qcut 1.6) needs
    a plan for its largest box: the engine makes it and the batch's rows
    equal the oracle's."""
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    Ls = [37.0, 38.2, 39.5, 40.7]
    boxes = [np.diag([L] * 3).astype(np.float32) for L in Ls]
    base, syms = O.synthetic_system(2000, 1.0, seed=91)
    frames = [(base * L).astype(np.float32) for L in Ls]
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((N_CELL,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, DQ, 1.6), boxes[0],
                     max_atoms=2000, max_frames=4)
    eng.forward_boxes(torch.from_numpy(np.stack(frames)), torch.from_numpy(ids), np.stack(boxes))
    torch.cuda.synchronize()
    assert eng.plans_created == 2
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    plan0 = O.make_plan(n3, boxes[0], 4, 2.0, None, DQ, 1.6)
    make = _plan_of(plan0)
    Ho = np.zeros(plan0.hp.num_bins)
    Co = np.zeros(plan0.hp.num_bins)
    for f, box in zip(frames, boxes):
        p = O.make_plan(n3, box, 4, 2.0, None, DQ, 1.6)
        p.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, p.hp.spacing)
        h, c = O.frame_histogram(p, f, ids, symbols, box)
        Ho += h
        Co += c
    assert np.array_equal(C, Co)
    sel = (np.arange(len(C)) >= 1) & (Co > 0)
    rel = np.abs(H[sel] / C[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    assert rel.max() <= 1e-5


def test_exact_bin_edges_in_device_tables(gpu):
    """Boxes whose shells sit exactly on bin edges (a frame with the first
    frame's box, whose spacing is the reset dq; one with exactly twice the
    box, whose shells |m| = 2b land on edges b; and a box scaled by 2^-3):
    the device tables' exact decisions give the host tables' counts."""
    import torch
    boxes = [np.diag([40.0] * 3), np.diag([80.0] * 3), np.diag([40.0] * 3), np.diag([35.0] * 3),
             np.diag([40.0 * 0.875] * 3)]
    boxes = [b.astype(np.float32) for b in boxes]
    base, syms = O.synthetic_system(1500, 1.0, seed=92)
    frames = [(base * b[0, 0]).astype(np.float32) for b in boxes]
    symbols, ids = O.element_ids(syms)
    xyz = torch.from_numpy(np.stack(frames))
    idt = torch.from_numpy(ids)
    a = _engine(symbols, boxes[0], len(boxes))
    a.forward_boxes(xyz, idt, np.stack(boxes))
    b = _engine(symbols, boxes[0], 1)
    for f, box in enumerate(boxes):
        b.forward(xyz[f:f + 1], idt, box=box)
    torch.cuda.synchronize()
    Ca, Cb = a.C.cpu().numpy(), b.C.cpu().numpy()
    Ha, Hb = a.H.cpu().numpy(), b.H.cpu().numpy()
    a.close()
    b.close()
    assert np.array_equal(Ca, Cb)
    sel = (np.arange(len(Ca)) >= 1) & (Cb > 0)
    assert np.max(np.abs(Ha[sel] - Hb[sel]) / np.abs(Hb[sel])) <= 1e-6
