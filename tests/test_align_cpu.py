"""Host-side checks of the alignment parameterisation (no GPU): the chain rule from dL/dH to dL/dp
and from a pyramid level's dL/dH_l to the base level, against float64 autograd of the same maps,
and the synthetic homographies' ranges."""
import math

import numpy as np
import torch

from mivisionx_amd import align


def test_identity_and_true_parameter_ranges():
    p0 = torch.zeros((1, 8), dtype=torch.float64)
    assert torch.allclose(align.H_from_p(p0, 1024, 1024)[0], torch.eye(3, dtype=torch.float64))
    p = align.true_params(64, 1024, 1024)
    H = align.H_from_p(torch.as_tensor(p), 1024, 1024)
    # the centre moves by the translation only; rotation angle from the normalised block
    c = torch.tensor([512.0, 512.0, 1.0], dtype=torch.float64)
    for i in range(64):
        th = math.atan2(p[i, 3], 1 + p[i, 0])
        assert abs(math.degrees(th)) <= 10.0 + 1e-9
        q = H[i] @ c
        q = q[:2] / q[2]
        assert abs(float(q[0]) - 512.0) <= 40.0 + 1e-6 and abs(float(q[1]) - 512.0) <= 40.0 + 1e-6
        assert abs(p[i, 6]) <= 2e-4 * 512 + 1e-12 and abs(p[i, 7]) <= 2e-4 * 512 + 1e-12


def test_grad_p_chain_rule_matches_autograd():
    torch.manual_seed(0)
    G = torch.randn((3, 9), dtype=torch.float64)
    p = (0.05 * torch.randn((3, 8), dtype=torch.float64)).requires_grad_(True)
    L = (align.H_from_p(p, 640, 480).reshape(3, 9) * G).sum()
    (want,) = torch.autograd.grad(L, p)
    got = align.grad_p_from_grad_H(G, 640, 480)
    assert torch.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_level_gradient_matches_autograd():
    torch.manual_seed(1)
    H = (torch.eye(3, dtype=torch.float64) + 0.01 * torch.randn((2, 3, 3), dtype=torch.float64)).requires_grad_(True)
    for level in range(3):
        s = 0.5 ** level
        S = torch.diag(torch.tensor([s, s, 1.0], dtype=torch.float64))
        Si = torch.diag(torch.tensor([1 / s, 1 / s, 1.0], dtype=torch.float64))
        G = torch.randn((2, 9), dtype=torch.float64)
        L = ((S @ H @ Si).reshape(2, 9) * G).sum()
        (want,) = torch.autograd.grad(L, H)
        got = align.level_grad_to_base(G, level)
        assert torch.allclose(got, want.reshape(2, 9), rtol=1e-12, atol=1e-12)
        assert torch.allclose(align.level_H(H.detach(), level).double(), (S @ H.detach() @ Si).reshape(2, 9).float().double())


def test_corner_error_is_zero_for_equal_and_exact_for_a_shift():
    H = torch.eye(3, dtype=torch.float64).repeat(2, 1, 1)
    T = H.clone()
    T[:, 0, 2] = 3.0
    T[:, 1, 2] = 4.0
    assert float(align.corner_error(H, H, 100, 80).max()) == 0.0
    assert np.allclose(align.corner_error(H, T, 100, 80).numpy(), 5.0)
