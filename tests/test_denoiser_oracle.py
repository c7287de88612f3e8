"""The denoiser's CPU oracle (oracle/denoiser_ref.py) pinned by mathematics:
its analytic gradient against central finite differences (float64, no
BF16 rounding), the identity div_b(curl_b A) = 0 that makes the correction
charge-conserving, and the adjoint identity of the curl."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import denoiser_ref as dr  # noqa: E402


def div_b(J):
    d = lambda f, ax: f - np.roll(f, 1, axis=ax)
    return d(J[..., 0], 2) + d(J[..., 1], 1) + d(J[..., 2], 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_curl_is_divergence_free_and_adjoint(dim):
    rng = np.random.default_rng(dim)
    shape = (4, 5, 6) if dim == 3 else (1, 7, 6)
    A = rng.standard_normal(shape + (3,))
    c = dr.curl_b(A, dim)
    assert np.abs(div_b(c)).max() < 1e-12
    g = rng.standard_normal(shape + (3,))
    lhs = np.sum(c * g)
    rhs = np.sum(A * dr.curl_b_adj(g, dim))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_curl_matches_yee_e_update_stencil():
    """the correction uses the E update's curl: Jx gains Az(j)-Az(j-1) - (Ay(k)-Ay(k-1))"""
    A = np.zeros((3, 3, 3, 3))
    A[1, 1, 1, 2] = 1.0     # Az at (x=1,y=1,z=1)
    c = dr.curl_b(A, 3)
    assert c[1, 1, 1, 0] == 1.0 and c[1, 2, 1, 0] == -1.0     # d/dy Az (index order z,y,x)
    assert c[1, 1, 1, 1] == -1.0 and c[1, 1, 2, 1] == 1.0     # -d/dx Az


@pytest.mark.parametrize("dim,lam", [(2, 0.0), (3, 0.0), (2, 2.0), (3, 2.0)])
def test_oracle_gradient_matches_finite_differences(dim, lam):
    rng = np.random.default_rng(10 + dim)
    shape = (3, 4, 5) if dim == 3 else (1, 6, 5)
    net = dr.Net(dim, width=4, nlayers=3)
    p = rng.standard_normal(net.nparams()) * 0.3
    net.set_flat(p)
    jlo = rng.standard_normal((3,) + shape)
    jt = rng.standard_normal((3,) + shape)
    loss, g, _ = net.loss_and_grad(jlo, jt, 1.3, bf=False, lam=lam)
    idx = rng.choice(p.size, 25, replace=False)
    eps = 1e-6
    for k in idx:
        q = p.copy(); q[k] += eps; net.set_flat(q)
        lp, _, _ = net.loss_and_grad(jlo, jt, 1.3, bf=False, grad=False, lam=lam)
        q[k] -= 2 * eps; net.set_flat(q)
        lm, _, _ = net.loss_and_grad(jlo, jt, 1.3, bf=False, grad=False, lam=lam)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.0e-5, -2.5], np.float32)
    r = dr.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0078125                      # rounds up
    assert r[4] == -2.5


def test_amplitude_term_is_the_projection_of_the_error():
    """lam <d,t>^2/(N<t,t>) is lam * N * mean(d_par^2) for d_par the part of
    the error along the target: a prediction that is the target scaled by
    (1 + e) costs (1 + lam) e^2 <t,t>/N"""
    rng = np.random.default_rng(5)
    net = dr.Net(2, width=4, nlayers=2)
    net.set_flat(np.zeros(net.nparams()))      # A = 0: prediction = input
    jt = rng.standard_normal((3, 1, 6, 5))
    e = 0.1
    for lam in (0.0, 3.0):
        loss, _, _ = net.loss_and_grad((1 + e) * jt, jt, 1.0, bf=False, grad=False, lam=lam)
        assert abs(loss - (1 + lam) * e * e * np.sum(jt * jt) / jt.size) < 1e-12
