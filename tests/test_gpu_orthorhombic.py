"""Orthorhombic cells with unequal reciprocal spacings (L_a nn_a / n_a not all
equal): the engine's point-keyed q tables against the FP64 oracle, whose
bins are pinned by a literal enumeration of the reference's loop and whose
spectrum by the direct structure factor (tests/test_oracle_pins.py).
Bars: counts bit-exact, I(q) within 1e-5 relative, gradient rel-L2 <= 2e-4."""
import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _positions(n_atoms, Ls, seed):
    pos, syms = O.synthetic_system(n_atoms, 1.0, seed=seed)     # uniform in [0, 1)
    return (pos * np.asarray(Ls)).astype(np.float32), syms


def _engine(n3, Ls, symbols, qcut, order=4, frames=1, n_atoms=1, flags=0, tile=(0, 0)):
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n3, nn = resolve_grid(n3, 2.0)
    spec = SystemSpec(n3, nn, order, tuple(symbols), 2.0, 0.0, qcut)
    return SaxsEngine(spec, np.diag(Ls), max_atoms=n_atoms, max_frames=frames, flags=flags,
                      tile=tile)


def _compare(H, C, Ho, Co, label):
    assert np.array_equal(C, Co), f"[{label}] counts {C} vs {Co}"
    sel = (np.arange(len(C)) >= 1) & (Co > 0)
    rel = np.abs(H[sel] / C[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    print(f"[{label}] max rel err {rel.max():.3e} over {sel.sum()} bins", flush=True)
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("n3,Ls,n_atoms,qcut,flags", [
    ((32, 32, 32), (40.0, 50.0, 60.0), 2000, 100.0, 0),   # full spectrum
    ((64, 64, 64), (60.0, 70.0, 80.0), 5000, 0.6, 0),     # pruned z, y and x passes
    ((64, 64, 64), (60.0, 70.0, 80.0), 5000, 0.6, 1),     # per-element spreading
    ((64, 64, 64), (60.0, 70.0, 80.0), 5000, 0.6, 2),     # full transforms
    ((30, 28, 36), (33.0, 40.0, 37.0), 2500, 1.2, 0),     # mixed radix, unequal n
    ((64, 48, 64), (80.0, 45.0, 64.0), 4000, 0.7, 0),     # pruned x/z, mixed-radix y
])
def test_orthorhombic_forward(gpu, n3, Ls, n_atoms, qcut, flags):
    import torch
    pos, syms = _positions(n_atoms, Ls, seed=31)
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, Ls, symbols, qcut, n_atoms=n_atoms, flags=flags)
    eng.forward(torch.from_numpy(pos[None]), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    print(f"kept {eng.kept()}", flush=True)
    eng.close()
    plan = O.make_plan(n3, np.diag(Ls), 4, 2.0, None, 0.0, qcut)
    assert not plan.isotropic
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, np.diag(Ls))
    _compare(H, C, Ho, Co, f"ortho {n3} {Ls} flags={flags}")


def test_orthorhombic_plain_path(gpu):
    import torch
    n3, Ls = (32, 32, 32), (36.0, 44.0, 52.0)
    pos, syms = _positions(2000, Ls, seed=32)
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, Ls, symbols, 0.9, n_atoms=2000)
    eng.forward_plain(torch.from_numpy(pos[None]), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    plan = O.make_plan(n3, np.diag(Ls), 4, 2.0, None, 0.0, 0.9)
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, np.diag(Ls))
    _compare(H, C, Ho, Co, "ortho plain")


def test_box_change_switches_table_keys(gpu):
    """A cubic first frame (shell keys), then an orthorhombic box (point keys),
    then cubic again: each frame binned with its own box like the oracle."""
    import torch
    n3, L0 = (32, 32, 32), 40.0
    boxes = [(40.0, 40.0, 40.0), (40.0, 39.0, 38.5), (39.5, 39.5, 39.5)]
    pos, syms = _positions(2000, (L0, L0, L0), seed=33)
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, boxes[0], symbols, 0.8, n_atoms=2000)
    H_or = np.zeros(eng.hist.num_bins)
    C_or = np.zeros(eng.hist.num_bins)
    plan0 = O.make_plan(n3, np.diag(boxes[0]), 4, 2.0, None, 0.0, 0.8)
    for Ls in boxes:
        p = (pos * np.asarray(Ls) / L0).astype(np.float32)
        eng.forward(torch.from_numpy(p[None]), torch.from_numpy(ids), box=np.diag(Ls))
        plan = O.make_plan(n3, np.diag(Ls), 4, 2.0, None, 0.0, 0.8)
        # dq, qcut and the bin count stay those of the first frame (reference NPT)
        plan.hp = O.HistParams(plan0.hp.dq, plan0.hp.qcut, plan0.hp.num_bins, plan.hp.spacing)
        h, c = O.frame_histogram(plan, p, ids, symbols, np.diag(Ls))
        H_or += h
        C_or += c
    torch.cuda.synchronize()
    _compare(eng.H.cpu().numpy(), eng.C.cpu().numpy(), H_or, C_or, "box changes")
    eng.close()


@pytest.mark.parametrize("n3,Ls,qcut,order", [((32, 32, 32), (40.0, 50.0, 60.0), 0.9, 4),
                                              ((64, 64, 64), (60.0, 70.0, 80.0), 0.6, 4),
                                              ((64, 64, 64), (60.0, 70.0, 80.0), 0.6, 6)])
def test_orthorhombic_adjoint(gpu, n3, Ls, qcut, order):
    import torch
    pos, syms = _positions(3000, Ls, seed=34)
    fr = [pos, (pos + np.random.default_rng(1).normal(0, 0.3, pos.shape)).astype(np.float32)]
    symbols, ids = O.element_ids(syms)
    eng = _engine(n3, Ls, symbols, qcut, order=order, frames=2, n_atoms=3000)
    xyz = torch.from_numpy(np.stack(fr))
    idt = torch.from_numpy(ids)
    Hf, _ = eng.forward_frames(xyz, idt)
    rng = np.random.default_rng(7)
    gH = rng.normal(size=tuple(Hf.shape)) / np.maximum(np.abs(Hf.cpu().numpy()), 1.0)
    grad = eng.backward(xyz, idt, torch.from_numpy(gH)).cpu().numpy()
    eng.close()
    plan = O.make_plan(n3, np.diag(Ls), order, 2.0, None, 0.0, qcut)
    for f in range(2):
        go = O.grad_linear_H(plan, fr[f], ids, symbols, np.diag(Ls), gH[f])
        rel = np.linalg.norm(grad[f] - go) / np.linalg.norm(go)
        mx = np.abs(grad[f] - go).max() / np.abs(go).max()
        print(f"[ortho adjoint {n3} P={order} frame {f}] rel-L2 {rel:.2e} max {mx:.2e}", flush=True)
        assert rel <= 2e-4 and mx <= 1e-3
