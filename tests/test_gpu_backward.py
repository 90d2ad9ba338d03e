"""Backward of the k-NN distances (needs a GPU).

Checked against (a) torch autograd through a dense FP64 distance matrix on
the CPU with the same indices, (b) FP64 central finite differences with the
indices fixed, and (c) the closed form on the 65,536-point sphere workload.
"""
import numpy as np
import pytest

import oracle
from pdist_aggregation_amd.autograd import knn_distances

pytestmark = pytest.mark.gpu


def sphere(rs, n):
    """Unit-sphere points with N(0, 0.01^2) radial jitter."""
    u = rs.randn(n, 3)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return u * (1.0 + 0.01 * rs.randn(n, 1))


def _dense_reference(torch, Y, X, idx, w, squared):
    Yt = torch.tensor(Y, dtype=torch.float64, requires_grad=True)
    Xt = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    D = ((Xt[:, None, :] - Yt[None, :, :]) ** 2).sum(-1)
    Dk = D.gather(1, torch.as_tensor(idx, dtype=torch.int64))
    if not squared:
        Dk = Dk.sqrt()
    (torch.as_tensor(w) * Dk).sum().backward()
    return Yt.grad.numpy(), Xt.grad.numpy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("squared", [True, False])
@pytest.mark.parametrize("d", [3, 16, 100, 300])
def test_backward_matches_dense_autograd(require_gpu, dtype, squared, d):
    torch = require_gpu
    rs = np.random.RandomState(d)
    Y = rs.rand(512, d).astype(dtype)
    X = rs.rand(384, d).astype(dtype)
    k = 16
    Yd = torch.tensor(Y, device="cuda", requires_grad=True)
    Xd = torch.tensor(X, device="cuda", requires_grad=True)
    dist, idx = knn_distances(Yd, Xd, k, squared=squared)
    w = rs.randn(X.shape[0], k)
    (dist * torch.tensor(w, dtype=dist.dtype, device="cuda")).sum().backward()
    gy_ref, gx_ref = _dense_reference(torch, Y, X, idx.cpu().numpy(), w, squared)
    tol = 2e-4 if dtype == np.float32 else 1e-10
    scale_x = np.abs(gx_ref).max()
    scale_y = np.abs(gy_ref).max()
    np.testing.assert_allclose(Xd.grad.cpu().numpy(), gx_ref, rtol=tol, atol=tol * scale_x)
    np.testing.assert_allclose(Yd.grad.cpu().numpy(), gy_ref, rtol=tol, atol=tol * scale_y)


def test_backward_finite_differences_fp64(require_gpu):
    """512 x 512 sphere subset, k=16, loss = sum of squared k-NN distances;
    FP64 central differences with step 1e-6 at fixed indices."""
    torch = require_gpu
    rs = np.random.RandomState(0)
    X = sphere(rs, 512)
    Y = sphere(rs, 512)
    Yd = torch.tensor(Y, device="cuda", requires_grad=True)
    Xd = torch.tensor(X, device="cuda", requires_grad=True)
    dist, idx = knn_distances(Yd, Xd, 16, squared=True)
    dist.sum().backward()
    idx = idx.cpu().numpy()
    gx, gy = Xd.grad.cpu().numpy(), Yd.grad.cpu().numpy()

    def loss(Xv, Yv):
        diff = Xv[:, None, :] - Yv[idx]
        return float((diff ** 2).sum())

    h = 1e-6
    for _ in range(40):
        which = rs.randint(2)
        i, f = rs.randint(512), rs.randint(3)
        A = (X if which == 0 else Y).copy()
        A[i, f] += h
        lp = loss(A, Y) if which == 0 else loss(X, A)
        A[i, f] -= 2 * h
        lm = loss(A, Y) if which == 0 else loss(X, A)
        fd = (lp - lm) / (2 * h)
        an = gx[i, f] if which == 0 else gy[i, f]
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (which, i, f, fd, an)


@pytest.mark.slow
def test_backward_sphere_workload(require_gpu):
    """65,536 x 65,536 points on the unit sphere, k=16, float32: forward +
    backward; gradients against the closed form in FP64."""
    torch = require_gpu
    rs = np.random.RandomState(0)
    X = sphere(rs, 65536).astype(np.float32)
    Y = sphere(rs, 65536).astype(np.float32)
    Yd = torch.tensor(Y, device="cuda", requires_grad=True)
    Xd = torch.tensor(X, device="cuda", requires_grad=True)
    dist, idx = knn_distances(Yd, Xd, 16, squared=True)
    dist.sum().backward()
    idx = idx.cpu().numpy()
    # forward parity on a subsample
    sub = np.arange(0, 65536, 64)
    rd, ri, rb = oracle.exact_knn_f64(X[sub], Y, 16, return_boundary=True)
    bad, _, _ = oracle.near_tie_report(X[sub], Y, idx[sub], rd, ri, rb, 2e-5)
    assert bad == 0
    diff = X.astype(np.float64)[:, None, :] - Y.astype(np.float64)[idx]
    gx_ref = 2.0 * diff.sum(1)
    gy_ref = np.zeros(Y.shape)
    np.add.at(gy_ref, idx.ravel(), -2.0 * diff.reshape(-1, 3))
    np.testing.assert_allclose(Xd.grad.cpu().numpy(), gx_ref, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(Yd.grad.cpu().numpy(), gy_ref, rtol=1e-4, atol=1e-4)
