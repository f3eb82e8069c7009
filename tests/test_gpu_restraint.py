"""The differentiable restraint configuration at reduced size: order-6
splines, bins as requested (exact dq policy, 100 bins from 0.005 to 0.5), a
chi^2 loss against the profile of a structure displaced by N(0, 1 A^2) with
1 % synthetic errors.

* the FP64 oracle gradient of chi^2 is pinned by central differences
  (h = 1e-3 A) along random directions over 200 random atoms, computed on the
  oracle's float64 forward;
* the engine's float32 gradient (autograd through pycusaxs_amd.autograd)
  matches the oracle gradient to 2e-4 relative L2.
"""
import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

L, N_ATOMS, N_CELL = 40.0, 4000, 32
DQ, QCUT = 0.005, 0.5


def setup():
    box = np.diag([L] * 3)
    base, syms = O.synthetic_system(N_ATOMS, L, seed=4)
    target_pos = O.frame_positions(base, 4, 1, 1.0)
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((N_CELL,) * 3, box, 6, 2.0, None, DQ, QCUT, "exact")
    return box, base, target_pos, symbols, ids, plan


def oracle_chi2_and_grad(plan, pos, ids, symbols, box, T, s, sel):
    H, C = O.frame_histogram(plan, pos, ids, symbols, box, literal=False)
    I = np.where(C > 0, H / np.where(C > 0, C, 1), 0.0)
    r = np.where(sel, (I - T) / s, 0.0)
    loss = float((r ** 2).sum())
    gI = np.where(sel, 2 * r / s, 0.0)
    gH = np.where(C > 0, gI / np.where(C > 0, C, 1), 0.0)
    return loss, O.grad_linear_H(plan, pos, ids, symbols, box, gH)


def test_restraint_gradient(gpu):
    import torch
    from pycusaxs_amd.autograd import chi2, saxs_profile
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    box, base, target_pos, symbols, ids, plan = setup()
    assert plan.hp.num_bins == 101
    Ht, Ct = O.frame_histogram(plan, target_pos, ids, symbols, box, literal=False)
    sel = (Ct > 0) & (np.arange(len(Ct)) >= 1)
    T = np.where(Ct > 0, Ht / np.where(Ct > 0, Ct, 1), 0.0)
    s = 0.01 * np.abs(T) + 1e-30
    # oracle gradient, pinned by directional central differences (float64 coords)
    O.EXACT_COORDS = True
    try:
        pos64 = base.astype(np.float64)
        loss0, g64 = oracle_chi2_and_grad(plan, pos64, ids, symbols, box, T, s, sel)
        rng = np.random.default_rng(0)
        atoms = rng.choice(N_ATOMS, size=200, replace=False)
        h = 1e-3
        for d in range(3):
            v = np.zeros_like(pos64)
            v[atoms] = rng.normal(size=(200, 3))
            lp = oracle_chi2_and_grad(plan, pos64 + h * v, ids, symbols, box, T, s, sel)[0]
            lm = oracle_chi2_and_grad(plan, pos64 - h * v, ids, symbols, box, T, s, sel)[0]
            fd = (lp - lm) / (2 * h)
            an = float((g64 * v).sum())
            print(f"direction {d}: FD {fd:.8e} analytic {an:.8e}", flush=True)
            assert abs(fd - an) <= 1e-5 * abs(an) + 1e-9 * abs(loss0)
    finally:
        O.EXACT_COORDS = False
    # oracle gradient with the reference's float32 coordinates vs the engine
    _, go = oracle_chi2_and_grad(plan, base, ids, symbols, box, T, s, sel)
    n3, nn = resolve_grid((N_CELL,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 6, tuple(symbols), 2.0, DQ, QCUT, "exact"), box,
                     N_ATOMS, 1)
    x = torch.from_numpy(base).cuda()[None].clone().requires_grad_(True)
    idt = torch.from_numpy(ids).cuda()
    I, _ = saxs_profile(x, eng, idt)
    Tt = torch.from_numpy(T).cuda()
    st = torch.from_numpy(s).cuda()
    loss = chi2(I[0], Tt, st, torch.from_numpy(sel).cuda())
    loss.backward()
    g = x.grad[0].cpu().numpy()
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"chi2 engine {loss.item():.6e} oracle {oracle_chi2_and_grad(plan, base, ids, symbols, box, T, s, sel)[0]:.6e}; grad rel-L2 {rel:.2e}", flush=True)
    assert rel <= 2e-4
    eng.close()


def test_restraint_step_graph_replays(gpu):
    """The captured restraint step (autograd.RestraintStep) against the
    eager autograd step on three coordinate sets, with eager engine work on
    the same plan and synchronisation between the replays: each replay's
    loss and gradient equal the eager step's."""
    import torch
    from pycusaxs_amd.autograd import RestraintStep, chi2, saxs_profile
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    box, base, target_pos, symbols, ids, _ = setup()
    n3, nn = resolve_grid((N_CELL,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 6, tuple(symbols), 2.0, DQ, QCUT, "exact"), box,
                     max_atoms=len(base), max_frames=1)
    idt = torch.from_numpy(ids).cuda()
    with torch.no_grad():
        T, C = saxs_profile(torch.from_numpy(target_pos.astype(np.float32)).cuda()[None], eng, idt)
    sel = (C[0] > 0) & (torch.arange(C.shape[1], device=C.device) >= 1)
    s = 0.01 * T[0].abs() + 1e-30

    def loss_fn(I):
        return chi2(I[0], T[0], s, sel)

    def eager(x):
        xg = x.clone().requires_grad_(True)
        I, _ = saxs_profile(xg, eng, idt)
        loss = loss_fn(I)
        loss.backward()
        return loss.detach(), xg.grad

    xs = [torch.from_numpy(O.frame_positions(base, 4, 10 + k, 0.3).astype(np.float32)).cuda()[None]
          for k in range(3)]
    step = RestraintStep(eng, idt, xs[0], loss_fn)
    for k, x in enumerate(xs):
        l_ref, g_ref = eager(x)                     # eager work on the same plan
        torch.cuda.synchronize()
        loss, grad = step(x)
        torch.cuda.synchronize()
        rel = ((grad - g_ref).norm() / g_ref.norm()).item()
        print(f"[replay {k}] loss {loss.item():.6e} (eager {l_ref.item():.6e}), grad rel-L2 {rel:.2e}",
              flush=True)
        assert torch.allclose(loss, l_ref, rtol=1e-9, atol=0)
        assert rel <= 1e-6
    eng.close()
