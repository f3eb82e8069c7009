"""Triclinic cells (rows a, b, c of a lower-triangular box, as MDAnalysis gives
them): fractional coordinates through the true inverse u = M x, q = sum m_a b_a
with signed point-keyed tables, each half-spectrum point binned at its own
|q|.  The engine against the FP64 oracle, whose triclinic path is pinned by
the direct structure factor of on-node atoms, a literal count enumeration
and finite differences (tests/test_oracle_*.py).  Bars: counts bit-exact,
I(q) within 1e-5 relative, gradient rel-L2 <= 2e-4."""
import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5
BOX_A = np.array([[30.0, 0, 0], [7.5, 28.0, 0], [-5.0, 6.25, 33.0]])
BOX_B = np.array([[60.0, 0, 0], [15.0, 55.0, 0], [-10.0, 12.0, 66.0]])
BOX_TO = np.array([[61.3, 0, 0], [-20.4, 57.7, 0], [-20.4, -28.9, 50.0]])   # octahedron-like


def _positions(n_atoms, B, seed):
    u, syms = O.synthetic_system(n_atoms, 1.0, seed=seed)             # fractions in [0, 1)
    return (u @ B).astype(np.float32), syms


def _engine(n3, B, symbols, qcut, order=4, frames=1, n_atoms=1, flags=0):
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n3, nn = resolve_grid(n3, 2.0)
    spec = SystemSpec(n3, nn, order, tuple(symbols), 2.0, 0.0, qcut)
    return SaxsEngine(spec, B, max_atoms=n_atoms, max_frames=frames, flags=flags)


def _compare(H, C, Ho, Co, label):
    assert np.array_equal(C, Co), f"[{label}] counts {C} vs {Co}"
    sel = (np.arange(len(C)) >= 1) & (Co > 0)
    rel = np.abs(H[sel] / C[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    print(f"[{label}] max rel err {rel.max():.3e} over {sel.sum()} bins", flush=True)
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("n3,B,n_atoms,qcut,flags", [
    ((32, 32, 32), BOX_A, 2000, 100.0, 0),      # full spectrum
    ((32, 32, 32), BOX_A, 2000, 0.9, 0),
    ((64, 64, 64), BOX_B, 5000, 0.5, 0),        # pruned passes
    ((64, 64, 64), BOX_B, 5000, 0.5, 1),        # per-element spreading
    ((64, 64, 64), BOX_B, 5000, 0.5, 2),        # full transforms
    ((48, 45, 40), BOX_TO, 4000, 0.6, 0),       # mixed radix, unequal grid
])
def test_triclinic_forward(gpu, n3, B, n_atoms, qcut, flags):
    import torch
    pos, syms = _positions(n_atoms, B, seed=41)
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, B, symbols, qcut, n_atoms=n_atoms, flags=flags)
    eng.forward(torch.from_numpy(pos[None]), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    print(f"kept {eng.kept()}", flush=True)
    eng.close()
    plan = O.make_plan(n3, B, 4, 2.0, None, 0.0, qcut)
    assert plan.triclinic
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, B)
    _compare(H, C, Ho, Co, f"tric {n3} flags={flags}")


def test_triclinic_plain_path(gpu):
    import torch
    n3 = (32, 32, 32)
    pos, syms = _positions(2000, BOX_A, seed=42)
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, BOX_A, symbols, 0.9, n_atoms=2000)
    eng.forward_plain(torch.from_numpy(pos[None]), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    plan = O.make_plan(n3, BOX_A, 4, 2.0, None, 0.0, 0.9)
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, BOX_A)
    _compare(H, C, Ho, Co, "tric plain")


def test_box_change_orthorhombic_triclinic(gpu):
    """Cubic first frame, then a sheared cell, then cubic again: each frame
    with its own cell (dq, qcut and bins of the first frame)."""
    import torch
    n3 = (32, 32, 32)
    boxes = [np.diag([40.0] * 3), np.array([[40.0, 0, 0], [4.0, 39.0, 0], [-2.0, 3.0, 40.0]]),
             np.diag([39.5] * 3)]
    u, syms = O.synthetic_system(2000, 1.0, seed=43)
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, boxes[0], symbols, 0.8, n_atoms=2000)
    plan0 = O.make_plan(n3, boxes[0], 4, 2.0, None, 0.0, 0.8)
    H_or = np.zeros(eng.hist.num_bins)
    C_or = np.zeros(eng.hist.num_bins)
    for B in boxes:
        p = (u @ B).astype(np.float32)
        eng.forward(torch.from_numpy(p[None]), torch.from_numpy(ids), box=B)
        plan = O.make_plan(n3, B, 4, 2.0, None, 0.0, 0.8)
        plan.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, plan.hp.spacing)
        h, c = O.frame_histogram(plan, p, ids, symbols, B)
        H_or += h
        C_or += c
    torch.cuda.synchronize()
    _compare(eng.H.cpu().numpy(), eng.C.cpu().numpy(), H_or, C_or, "ortho/tric box changes")
    eng.close()


@pytest.mark.parametrize("n3,B,qcut,order", [((32, 32, 32), BOX_A, 100.0, 4),
                                             ((32, 32, 32), BOX_A, 0.9, 6),
                                             ((64, 64, 64), BOX_B, 0.5, 4)])
def test_triclinic_adjoint(gpu, n3, B, qcut, order):
    import torch
    pos, syms = _positions(3000, B, seed=44)
    fr = [pos, (pos + np.random.default_rng(2).normal(0, 0.3, pos.shape)).astype(np.float32)]
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, B, symbols, qcut, order=order, frames=2, n_atoms=3000)
    xyz = torch.from_numpy(np.stack(fr))
    idt = torch.from_numpy(ids)
    Hf, _ = eng.forward_frames(xyz, idt)
    rng = np.random.default_rng(9)
    gH = rng.normal(size=tuple(Hf.shape)) / np.maximum(np.abs(Hf.cpu().numpy()), 1.0)
    grad = eng.backward(xyz, idt, torch.from_numpy(gH)).cpu().numpy()
    eng.close()
    plan = O.make_plan(n3, B, order, 2.0, None, 0.0, qcut)
    for f in range(2):
        go = O.grad_linear_H(plan, fr[f], ids, symbols, B, gH[f])
        rel = np.linalg.norm(grad[f] - go) / np.linalg.norm(go)
        mx = np.abs(grad[f] - go).max() / np.abs(go).max()
        print(f"[tric adjoint {n3} P={order} frame {f}] rel-L2 {rel:.2e} max {mx:.2e}", flush=True)
        assert rel <= 2e-4 and mx <= 1e-3
