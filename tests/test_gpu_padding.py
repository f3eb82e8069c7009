"""GPU parity of the bulk-solvent "given" pad mode (padding::given: a constant
per element, here the explicit-water values of AtomCounter) against the FP64
oracle, forward and adjoint.  Bars as in the other parity tests: counts
bit-exact, I(q) within 1e-5 relative, gradient rel-L2 <= 2e-4.

The given pad leaves a box-shaped term (rho_border - pad) * FFT(1_cell) in the
spectrum, unlike the border mean, so these cases also exercise the exact
(64-bit) fixed-point subtraction when the pad exceeds 2^31 units."""
import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _engine(n, L, symbols, order, qcut, frames, flags=0, tile=(0, 0), pads=None, dq=0.0):
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n3, nn = resolve_grid((n,) * 3, 2.0)
    spec = SystemSpec(n3, nn, order, tuple(symbols), 2.0, dq, qcut)
    return SaxsEngine(spec, np.diag([L] * 3), max_atoms=frames[0].shape[0],
                      max_frames=len(frames), flags=flags, tile=tile, pad_given=pads)


def _compare(H, C, Ho, Co, label):
    assert np.array_equal(C, Co), f"[{label}] bin counts differ"
    sel = (np.arange(len(C)) >= 1) & (Co > 0)
    rel = np.abs(H[sel] / C[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    print(f"[{label}] max rel err {rel.max():.3e} over {sel.sum()} bins", flush=True)
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("n,L,n_atoms,qcut,flags", [
    (64, 37.0, 5000, 0.5, 0),        # 5k atoms, 37 A, 64^3 -> 128^3, SPCE water
    (64, 37.0, 5000, 0.5, 1),        # per-element spreading workgroups
    (64, 37.0, 5000, 0.5, 2),        # full (unpruned) z/y transforms
    (64, 128.0, 4000, 0.3, 0),       # 8 A^3 voxels: H pad above 2^31 fixed-point units
    (24, 30.0, 1500, 100.0, 0),      # mixed radix, full spectrum
])
def test_given_padding_forward(gpu, n, L, n_atoms, qcut, flags):
    import torch
    from pycusaxs_amd.geometry import explicit_solvent_pads
    pos, syms = O.synthetic_system(n_atoms, L, seed=21)
    symbols, ids = O.element_ids(syms)
    pads = explicit_solvent_pads((L, L, L), (n, n, n), "SPCE", {"Na": 5, "Cl": 5})
    eng = _engine(n, L, symbols, 4, qcut, [pos], flags=flags, pads=pads)
    eng.forward(torch.from_numpy(pos[None].astype(np.float32)), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    box = np.diag([L] * 3)
    plan = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, qcut,
                       pad_given=O.solvent_pads((L, L, L), (n, n, n), "SPCE", {"Na": 5, "Cl": 5}))
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, box)
    _compare(H, C, Ho, Co, f"given n={n} L={L} flags={flags}")
    # switching back to the border mean reproduces the avg-mode oracle
    eng.set_padding(None)
    eng.reset()
    eng.forward(torch.from_numpy(pos[None].astype(np.float32)), torch.from_numpy(ids))
    torch.cuda.synchronize()
    plan_avg = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, qcut)
    Ha, Ca = O.frame_histogram(plan_avg, pos, ids, symbols, box)
    _compare(eng.H.cpu().numpy(), eng.C.cpu().numpy(), Ha, Ca, "avg after given")
    eng.close()


def test_given_padding_plain_path(gpu):
    import torch
    from pycusaxs_amd.geometry import explicit_solvent_pads
    n, L = 32, 30.0
    pos, syms = O.synthetic_system(2000, L, seed=22)
    symbols, ids = O.element_ids(syms)
    pads = explicit_solvent_pads((L, L, L), (n, n, n), "TIP3P", {})
    eng = _engine(n, L, symbols, 4, 0.8, [pos], pads=pads)
    eng.forward_plain(torch.from_numpy(pos[None].astype(np.float32)), torch.from_numpy(ids))
    torch.cuda.synchronize()
    box = np.diag([L] * 3)
    plan = O.make_plan((n,) * 3, box, 4, 2.0, None, 0.0, 0.8, pad_given=pads)
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, box)
    _compare(eng.H.cpu().numpy(), eng.C.cpu().numpy(), Ho, Co, "given plain")
    eng.close()


@pytest.mark.parametrize("n,L,n_atoms,qcut,order", [(32, 40.0, 3000, 0.8, 4),
                                                    (64, 40.0, 4000, 0.6, 4),
                                                    (64, 40.0, 4000, 0.6, 6)])
def test_given_padding_adjoint(gpu, n, L, n_atoms, qcut, order):
    """No pad-mean term in the adjoint: the engine gradient matches the
    oracle's given-mode gradient (itself pinned by finite differences)."""
    import torch
    base, syms = O.synthetic_system(n_atoms, L, seed=23)
    fr = [O.frame_positions(base, 23, f, 0.3) for f in range(2)]
    symbols, ids = O.element_ids(syms)
    pads = O.solvent_pads((L, L, L), (n, n, n), "SPCE", {})
    eng = _engine(n, L, symbols, order, qcut, fr, pads=pads)
    xyz = torch.from_numpy(np.stack(fr))
    idt = torch.from_numpy(ids)
    Hf, _ = eng.forward_frames(xyz, idt)
    rng = np.random.default_rng(5)
    gH = rng.normal(size=tuple(Hf.shape)) / np.maximum(np.abs(Hf.cpu().numpy()), 1.0)
    grad = eng.backward(xyz, idt, torch.from_numpy(gH)).cpu().numpy()
    eng.close()
    box = np.diag([L] * 3)
    plan = O.make_plan((n,) * 3, box, order, 2.0, None, 0.0, qcut, pad_given=pads)
    for f in range(len(fr)):
        go = O.grad_linear_H(plan, fr[f], ids, symbols, box, gH[f])
        rel = np.linalg.norm(grad[f] - go) / np.linalg.norm(go)
        mx = np.abs(grad[f] - go).max() / np.abs(go).max()
        print(f"[given adjoint n={n} P={order} frame {f}] rel-L2 {rel:.2e} max {mx:.2e}", flush=True)
        assert rel <= 2e-4 and mx <= 1e-3
