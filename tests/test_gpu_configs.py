"""The benchmark workloads at their own sizes, through the engine on the GPU:

  trajectory  100,000 atoms, 100 A, 128^3 -> 256^3, q <= 0.6 (3 of its frames)
              against the FP64 oracle: counts exact, I(q) within 1e-5
  restraint   50,000 atoms, 80 A, 128^3, order 6, chi^2 over 100 exact bins:
              the autograd gradient against the FP64 oracle gradient
  uneven      the 32 systems in one hipGraph (three replays) against the
              eager chains; the smallest, median and largest against the oracle
  large       1,000,000 atoms, 215 A, 9 elements, 256^3 -> 512^3, one frame:
              output-pruned transforms and adjoints against the full Stockham
              ones (the 512-point kernels are pinned against mathematics in
              test_gpu_parity.py)

Inputs come from pycusaxs_amd.workloads, the same seeded generators as
bench.py.  The oracle's transforms run on the host's usable CPUs.
"""
import os

import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5
C2_COUNTS = [34, 82, 178, 258, 450, 506, 786, 922, 1226, 1426, 1690, 2050, 2458, 2626, 3170,
             3402, 4042, 4362, 506]


@pytest.fixture
def threads():
    try:
        k = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        k = os.cpu_count() or 1
    used = O.set_fft_workers(min(16, k))
    yield used
    O.set_fft_workers(1)


def _rows_close(H, C, Ho, Co, tol, label):
    assert np.array_equal(C, Co), f"[{label}] counts differ: {C} vs {Co}"
    sel = (np.arange(len(C)) >= 1) & (Co > 0)
    rel = np.abs(H[sel] / C[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
    print(f"[{label}] max rel err {rel.max():.3e} over {sel.sum()} bins", flush=True)
    assert rel.max() <= tol


def test_trajectory_100k_128(gpu, threads):
    import torch
    from pycusaxs_amd import workloads as W
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    t = W.TRAJ
    xyz, symbols, ids, _ = W.trajectory(frames=3)
    box = np.diag([t["box"]] * 3)
    n3, nn = resolve_grid((t["grid"],) * 3, t["sigma"])
    spec = SystemSpec(n3, nn, t["order"], tuple(symbols), t["sigma"], t["dq"], t["qcut"])
    eng = SaxsEngine(spec, box, max_atoms=t["n_atoms"], max_frames=3)
    eng.forward(torch.from_numpy(xyz), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    assert list(C[1:20].astype(int) // 3) == C2_COUNTS
    plan = O.make_plan(n3, box, t["order"], t["sigma"], None, t["dq"], t["qcut"])
    Ho, Co = O.trajectory_histogram(plan, [(f, box) for f in xyz], ids, symbols)
    _rows_close(H, C, Ho, Co, REL_TOL, "trajectory 100k, 128^3, 3 frames")


def test_restraint_50k_order6_gradient(gpu, threads):
    import torch
    from pycusaxs_amd import workloads as W
    from pycusaxs_amd.autograd import chi2, saxs_profile
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    w = W.restraint()
    n3, nn = resolve_grid((w["n"],) * 3, 2.0)
    box = np.diag([w["L"]] * 3)
    eng = SaxsEngine(SystemSpec(n3, nn, w["order"], tuple(w["symbols"]), 2.0, w["dq"],
                                w["qcut"], "exact"), box, len(w["base"]), 1)
    idt = torch.from_numpy(w["ids"]).cuda()
    with torch.no_grad():
        T, Ct = saxs_profile(torch.from_numpy(w["target"]).cuda()[None], eng, idt)
    sel = (Ct[0] > 0) & (torch.arange(Ct.shape[1], device=Ct.device) >= 1)
    s = 0.01 * T[0].abs() + 1e-30
    x = torch.from_numpy(w["base"]).cuda()[None].clone().requires_grad_(True)
    I, Ce = saxs_profile(x, eng, idt)
    loss = chi2(I[0], T[0], s, sel)
    loss.backward()
    grad = x.grad[0].cpu().numpy()
    I_e = I[0].detach().cpu().numpy()
    eng.close()
    assert int(sel.sum()) > 60
    plan = O.make_plan(n3, box, w["order"], 2.0, None, w["dq"], w["qcut"], "exact")
    Ho, Co = O.frame_histogram(plan, w["base"], w["ids"], w["symbols"], box, literal=False)
    assert np.array_equal(Ce[0].cpu().numpy(), Co)
    Io = np.where(Co > 0, Ho / np.where(Co > 0, Co, 1), 0.0)
    m = sel.cpu().numpy()
    rel_I = np.abs(I_e[m] - Io[m]) / np.abs(Io[m])
    print(f"[restraint] I(q) max rel err {rel_I.max():.3e} over {m.sum()} bins", flush=True)
    assert rel_I.max() <= REL_TOL
    Tn, sn = T[0].cpu().numpy(), s.cpu().numpy()
    r = np.where(m, (Io - Tn) / sn, 0.0)
    gH = np.where(Co > 0, np.where(m, 2 * r / sn, 0.0) / np.where(Co > 0, Co, 1), 0.0)
    go = O.grad_linear_H(plan, w["base"], w["ids"], w["symbols"], box, gH)
    rel = np.linalg.norm(grad - go) / np.linalg.norm(go)
    mx = np.abs(grad - go).max() / np.abs(go).max()
    print(f"[restraint] gradient rel-L2 {rel:.3e}, max/|g|max {mx:.3e}", flush=True)
    assert rel <= 2e-4 and mx <= 1e-3


@pytest.mark.timeout(400)
def test_uneven_32_systems_graph(gpu, threads):
    import torch
    from pycusaxs_amd import workloads as W
    from pycusaxs_amd import _lib
    from pycusaxs_amd.batch import BatchedSystems, SystemInput
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    work = W.uneven_systems()
    systems = []
    for w in work:
        n3, nn = resolve_grid((w["n"],) * 3, 2.0)
        systems.append(SystemInput(SystemSpec(n3, nn, 4, tuple(w["symbols"]), 2.0, 0.005, 0.5),
                                   np.diag([w["L"]] * 3), w["n_atoms"], w["xyz"].shape[0]))
    bs = BatchedSystems(systems)
    for i, w in enumerate(work):
        bs.load(i, w["xyz"], w["ids"])
    bs.forward_eager()
    torch.cuda.synchronize()
    eager = [(h.clone(), c.clone()) for h, c in bs.histograms()]
    bs.capture()
    for rep in range(3):
        x0 = bs.xyz[0].clone()                  # eager work between the replays
        bs.xyz[0].copy_(x0)
        torch.cuda.synchronize()
        bs.forward()
        torch.cuda.synchronize()
        for i, (h, c) in enumerate(bs.histograms()):
            assert torch.equal(c, eager[i][1]), (rep, i)
            rel = ((h - eager[i][0]).abs() / eager[i][0].abs().clamp_min(1e-30))[1:]
            assert rel.max().item() < 1e-6, (rep, i, rel.max().item())
    print(f"[uneven] 3 replays of {len(work)} systems agree with the eager chains", flush=True)
    hist = [(h.cpu().numpy(), c.cpu().numpy()) for h, c in bs.histograms()]
    bs.close()
    order = sorted(range(len(work)), key=lambda i: work[i]["n_atoms"])
    small, median, large = order[0], order[len(order) // 2], order[-1]
    for i in (small, median):
        w = work[i]
        box = np.diag([w["L"]] * 3)
        plan = O.make_plan((w["n"],) * 3, box, 4, 2.0, None, 0.005, 0.5)
        Ho, Co = O.trajectory_histogram(plan, [(f, box) for f in w["xyz"]], w["ids"], w["symbols"])
        _rows_close(*hist[i], Ho, Co, REL_TOL, f"uneven system {i}: {w['n_atoms']} atoms, "
                                              f"{w['n']}^3, 8 frames")
    # the largest (300k atoms, 189^3 -> 378^3): its first frame alone, with the
    # plan the graph used, against the oracle
    w = work[large]
    box = np.diag([w["L"]] * 3)
    n3, nn = resolve_grid((w["n"],) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(w["symbols"]), 2.0, 0.005, 0.5), box,
                     max_atoms=w["n_atoms"], max_frames=1, flags=_lib.FLAG_FORWARD_ONLY)
    eng.forward(torch.from_numpy(w["xyz"][:1]), torch.from_numpy(w["ids"]))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    plan = O.make_plan(n3, box, 4, 2.0, None, 0.005, 0.5)
    Ho, Co = O.trajectory_histogram(plan, [(w["xyz"][0], box)], w["ids"], w["symbols"])
    _rows_close(H, C, Ho, Co, REL_TOL, f"uneven system {large}: {w['n_atoms']} atoms, "
                                       f"{w['n']}^3, frame 0")


def test_large_1M_512_pruned_matches_full_transforms(gpu):
    import torch
    from pycusaxs_amd import workloads as W
    from pycusaxs_amd import _lib
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    L, n = W.LARGE["box"], W.LARGE["grid"]
    base, symbols, ids = W.large_base()
    assert len(symbols) == 9
    xyz = torch.from_numpy(W.large_frame(base, 0)[None]).cuda()
    idt = torch.from_numpy(ids).cuda()
    n3, nn = resolve_grid((n,) * 3, 2.0)
    assert nn == (512, 512, 512)
    out = []
    for flags in (0, _lib.FLAG_NO_PRUNED_Z):
        eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, W.LARGE["qcut"]),
                         np.diag([L] * 3), max_atoms=len(ids), max_frames=1, tile=(4, 4),
                         flags=flags)
        Hf, Cf = eng.forward_frames(xyz, idt)
        gH = torch.from_numpy(np.random.default_rng(9).normal(size=tuple(Hf.shape))).cuda() \
            / Hf.abs().clamp_min(1.0)
        g = eng.backward(xyz, idt, gH)
        torch.cuda.synchronize()
        out.append((Hf[0].cpu().numpy(), Cf[0].cpu().numpy(), g[0].cpu().numpy(), eng.kept()))
        eng.close()
    (H1, C1, g1, k1), (H2, C2, g2, k2) = out
    assert k1 == k2 and k1[2] <= 512 // 8
    _rows_close(H1, C1, H2, C2, 2e-6, "1M atoms, 512^3: pruned vs full")
    rel = np.linalg.norm(g1 - g2) / np.linalg.norm(g2)
    print(f"[1M atoms, 512^3] adjoint pruned vs full rel-L2 {rel:.3e}", flush=True)
    assert rel <= 1e-5


def test_reference_cli_defaults_plan(gpu, threads):
    """The plan the reference CLI's defaults give (grid 128, scale 2.5 ->
    320^3 FFT, bin 0.01 reset to the grid spacing, qcut 2.0: kz = 80,
    ky = kx = 159, the full-length mixed-radix y and x transforms) on a
    sparse system, against the oracle."""
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    L = 100.0
    box = np.diag([L] * 3)
    pos, syms = O.synthetic_system(4000, L, seed=12, mix=(("C", 0.6), ("O", 0.3), ("N", 0.1)))
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((128,) * 3, 2.5)
    assert nn == (320, 320, 320)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.5, 0.01, 2.0), box,
                     max_atoms=len(pos), max_frames=1)
    eng.forward(torch.from_numpy(pos[None]), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    assert eng.kept() == (159, 159, 80)
    eng.close()
    plan = O.make_plan(n3, box, 4, 2.5, None, 0.01, 2.0)
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, box)
    _rows_close(H, C, Ho, Co, REL_TOL, "reference CLI defaults, 128 -> 320")
