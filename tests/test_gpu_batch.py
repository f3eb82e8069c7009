"""Batched uneven systems in one hipGraph (pycusaxs_amd.batch) and the
autotuner.  Every graph replay is checked against the eager launch chain
and the FP64 oracle, with eager work and synchronisation between replays."""
import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu


def make_systems():
    from pycusaxs_amd.batch import SystemInput
    from pycusaxs_amd.engine import SystemSpec, resolve_grid
    out, data = [], []
    for s, (n_atoms, L, n) in enumerate([(800, 22.0, 24), (1500, 27.0, 30), (600, 20.0, 21),
                                         (2500, 32.0, 32)]):
        base, syms = O.synthetic_system(n_atoms, L, seed=100 + s)
        frames = np.stack([O.frame_positions(base, 100 + s, f, 0.5) for f in range(3)])
        symbols, ids = O.element_ids(syms)
        n3, nn = resolve_grid((n,) * 3, 2.0)
        spec = SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, 1.2)
        out.append(SystemInput(spec, np.diag([L] * 3), n_atoms, 3))
        data.append((frames, symbols, ids, L, n))
    return out, data


def test_graph_replays_match_eager_and_oracle(gpu):
    import torch
    from pycusaxs_amd.batch import BatchedSystems
    systems, data = make_systems()
    bs = BatchedSystems(systems)
    for i, (frames, symbols, ids, L, n) in enumerate(data):
        bs.load(i, frames, ids)
    bs.forward_eager()
    torch.cuda.synchronize()
    eager = [(h.clone(), c.clone()) for h, c in bs.histograms()]
    bs.capture()
    for rep in range(3):
        # eager work between replays (a copy of the same inputs back in place)
        for i, (frames, symbols, ids, L, n) in enumerate(data):
            bs.load(i, frames, ids)
        torch.cuda.synchronize()
        bs.forward()
        torch.cuda.synchronize()
        for i, (h, c) in enumerate(bs.histograms()):
            assert torch.equal(c, eager[i][1]), (rep, i)
            rel = ((h - eager[i][0]).abs() / eager[i][0].abs().clamp_min(1e-30))[1:]
            assert rel.max().item() < 1e-6, (rep, i, rel.max().item())
    for i, (frames, symbols, ids, L, n) in enumerate(data):
        plan = O.make_plan((n,) * 3, np.diag([L] * 3), 4, 2.0, None, 0.0, 1.2)
        Ho, Co = O.trajectory_histogram(plan, [(f, np.diag([L] * 3)) for f in frames], ids, symbols)
        h, c = (t.cpu().numpy() for t in bs.histograms()[i])
        assert np.array_equal(c, Co)
        sel = (np.arange(len(c)) >= 1) & (Co > 0)
        rel = np.abs(h[sel] / c[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])
        print(f"system {i} (n={n}): max rel err {rel.max():.2e}", flush=True)
        assert rel.max() <= 1e-5
    bs.close()


def test_autotuner_picks_valid_shape_and_caches(gpu, tmp_path):
    from pycusaxs_amd.tuning import Autotuner
    from pycusaxs_amd.batch import BatchedSystems
    systems, data = make_systems()
    tuner = Autotuner(cache_path=tmp_path / "tune.json", frames=2, repeats=1)
    s = systems[3]
    tile, flags = tuner.tune(s.spec, s.n_atoms, s.box)
    assert tuner.lookup(s.spec, s.n_atoms, s.box) == (tile, flags)
    assert (tmp_path / "tune.json").exists()
    # a tuned plan computes the same histogram
    frames, symbols, ids, L, n = data[3]
    bs = BatchedSystems([s], tuner=tuner)
    bs.load(0, frames, ids)
    bs.forward()
    import torch
    torch.cuda.synchronize()
    plan = O.make_plan((n,) * 3, np.diag([L] * 3), 4, 2.0, None, 0.0, 1.2)
    Ho, Co = O.trajectory_histogram(plan, [(f, np.diag([L] * 3)) for f in frames], ids, symbols)
    h, c = (t.cpu().numpy() for t in bs.histograms()[0])
    assert np.array_equal(c, Co)
    sel = (np.arange(len(c)) >= 1) & (Co > 0)
    assert (np.abs(h[sel] / c[sel] - Ho[sel] / Co[sel]) / np.abs(Ho[sel] / Co[sel])).max() <= 1e-5
    bs.close()
