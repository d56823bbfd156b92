"""The torch binding's initial state from an energy-density field
(viral_amd.autograd.state_from_energy) under both EOSs: P(e) equals the
library's EOS, and its gradient dP/de (the s95p closed form's derivative
through the library) matches central differences of P(e)."""
import numpy as np
import pytest
import torch

from viral_amd import autograd, capi
from viral_amd.desc import make_desc, EOS_S95P, EOS_CONFORMAL, F_P, F_T00, F_EN, F_TEMP


@pytest.mark.parametrize("eos", [EOS_CONFORMAL, EOS_S95P])
def test_state_from_energy_pressure_and_gradient(eos):
    d = make_desc(4, 4, 3, eos=eos)
    e0 = np.random.default_rng(1).uniform(0.05, 60.0, d.cells())
    e = torch.tensor(e0, requires_grad=True)
    st = autograd.state_from_energy(d, e, 0.6)
    p, dpde, T, s = capi.eos_eval(d, e0)
    assert np.allclose(st[F_P].detach().numpy(), p, rtol=1e-15, atol=0)
    assert np.allclose(st[F_T00].detach().numpy(), e0, rtol=1e-15, atol=0)
    assert np.allclose(st[F_TEMP].detach().numpy(), T, rtol=1e-15, atol=0)
    w = torch.from_numpy(np.random.default_rng(2).standard_normal(d.cells()))
    (st[F_P] * w).sum().backward()
    h = 1e-6
    fd = (capi.eos_eval(d, e0 * (1 + h))[0] - capi.eos_eval(d, e0 * (1 - h))[0]) / (2 * h * e0)
    assert np.allclose(e.grad.numpy(), w.numpy() * fd, rtol=1e-7, atol=0)
