"""chi^2 restraint loss (pycusaxs_amd.autograd.chi2) on the CPU: a boolean
bin mask and the equivalent index list give the same loss and gradient, the
mask form keeps a finite zero gradient on unselected bins whose sigma is 0,
and the loss matches the closed form."""
import numpy as np
import torch

from pycusaxs_amd.autograd import chi2


def test_mask_and_index_forms_agree():
    rng = np.random.default_rng(5)
    I0 = torch.tensor(rng.random(40) + 0.5, dtype=torch.float64)
    T = torch.tensor(rng.random(40) + 0.5, dtype=torch.float64)
    s = torch.tensor(0.05 + 0.1 * rng.random(40), dtype=torch.float64)
    mask = torch.tensor(rng.random(40) < 0.6)
    mask[0] = False
    idx = torch.nonzero(mask).flatten()
    Ia = I0.clone().requires_grad_(True)
    Ib = I0.clone().requires_grad_(True)
    la, lb = chi2(Ia, T, s, mask), chi2(Ib, T, s, idx)
    la.backward()
    lb.backward()
    ref = (((I0 - T) / s) ** 2)[mask].sum()
    assert torch.allclose(la, ref, rtol=1e-14) and torch.allclose(lb, ref, rtol=1e-14)
    assert torch.allclose(Ia.grad, Ib.grad, rtol=1e-14, atol=0)
    assert torch.all(Ia.grad[~mask] == 0)
    assert torch.allclose(Ia.grad[mask], (2 * (I0 - T) / s ** 2)[mask], rtol=1e-14)


def test_mask_with_zero_sigma_outside_keeps_gradient_finite():
    I = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64, requires_grad=True)
    T = torch.tensor([1.5, 2.0, 2.0, 0.0], dtype=torch.float64)
    s = torch.tensor([0.5, 1.0, 0.0, 0.0], dtype=torch.float64)    # zero outside the mask
    mask = torch.tensor([True, True, False, False])
    loss = chi2(I, T, s, mask)
    loss.backward()
    assert loss.item() == 1.0
    assert torch.isfinite(I.grad).all()
    assert I.grad.tolist() == [-4.0, 0.0, 0.0, 0.0]
