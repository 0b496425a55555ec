"""GPU parity of the FP64 vector kernels and the symmetric eigensolver
(through the C ABI) against numpy."""
import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 1 << 20, 3_000_001])
def test_dot_nrm2(ctx, dev, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    dx, dy = dev.put(x), dev.put(y)
    assert ctx.dot(dx.data_ptr(), dy.data_ptr(), n) == pytest.approx(float(x @ y), rel=1e-12, abs=1e-12 * n)
    assert ctx.nrm2(dx.data_ptr(), n) == pytest.approx(float(np.linalg.norm(x)), rel=1e-13)


def test_multi_dot_lincomb_axpy_scal(ctx, dev):
    rng = np.random.default_rng(5)
    n = 123457
    V = [rng.standard_normal(n) for _ in range(5)]
    y = rng.standard_normal(n)
    dV = [dev.put(v) for v in V]
    dy = dev.put(y)
    got = ctx.multi_dot([t.data_ptr() for t in dV], dy.data_ptr(), n)
    ref = np.array([v @ y for v in V])
    assert relerr(got, ref) < 1e-12
    c = rng.standard_normal(5)
    out = dev.empty(n)
    ctx.lincomb([t.data_ptr() for t in dV], c, out.data_ptr(), n)
    assert relerr(dev.get(out, (n,)), sum(ci * v for ci, v in zip(c, V))) < 1e-13
    ctx.axpy(-0.75, dV[0].data_ptr(), dy.data_ptr(), n)
    assert relerr(dev.get(dy, (n,)), y - 0.75 * V[0]) < 1e-14
    ctx.scal(3.0, dy.data_ptr(), n)
    assert relerr(dev.get(dy, (n,)), 3 * (y - 0.75 * V[0])) < 1e-14


@pytest.mark.parametrize("dims,perm", [((3, 4, 5), (2, 0, 1)), ((7, 1, 9, 2), (3, 1, 0, 2)), ((50, 60), (1, 0))])
def test_permute_add(ctx, dev, dims, perm):
    rng = np.random.default_rng(9)
    # X has dims dimsX with dimsX[perm[j]] = dims[j]
    dimsX = [0] * len(dims)
    for j, p in enumerate(perm):
        dimsX[p] = dims[j]
    X = rng.standard_normal(dimsX)
    Y = rng.standard_normal(dims)
    ref = Y + 2.5 * np.transpose(X, perm)
    dX, dY = dev.put(X), dev.put(Y)
    ctx.permute(dX.data_ptr(), dY.data_ptr(), dims, perm, 2.5, 1.0)
    assert relerr(dev.get(dY, dims), ref) < 1e-15
    ctx.permute(dX.data_ptr(), dY.data_ptr(), dims, perm, 1.0, 0.0)
    assert relerr(dev.get(dY, dims), np.transpose(X, perm)) == 0.0


@pytest.mark.parametrize("n", [1, 2, 3, 10, 64, 65, 100, 128, 129, 400])
def test_syev(ctx, dev, n):
    """Single matrices of every size class (LDS-resident up to 64 and up to
    128, global above): eigenvalues vs numpy, orthonormal vectors,
    reconstruction, descending order."""
    rng = np.random.default_rng(1000 + n)
    X = rng.standard_normal((n, n))
    if n in (64, 128):
        # density-matrix-like: PSD with exact zero eigenvalues and degeneracies
        X = X[:, :20] @ X[:, :20].T
        X[:, :2] = 0
        X[:2, :] = 0
    else:
        X = X + X.T
    dA = dev.put(X)
    dw = dev.empty(n)
    ctx.syev(n, dA.data_ptr(), dw.data_ptr())
    w = dev.get(dw, (n,))
    U = dev.get(dA, (n, n))
    wref = np.linalg.eigvalsh(X)[::-1]
    scale = max(1.0, np.max(np.abs(wref)))
    assert np.max(np.abs(w - wref)) / scale < 1e-12
    assert np.all(np.diff(w) <= 0)  # descending
    assert np.max(np.abs(U.T @ U - np.eye(n))) < 1e-12
    assert np.max(np.abs(U @ np.diag(w) @ U.T - X)) / scale < 1e-12


def test_syev_batched_mixed_classes(ctx, dev):
    """One batched call over matrices of all three size classes, including
    degenerate and rank-deficient ones, stored back to back."""
    rng = np.random.default_rng(77)
    ns = [5, 64, 1, 90, 200, 33, 128, 0, 150]
    mats = []
    for k, n in enumerate(ns):
        X = rng.standard_normal((n, n))
        if k % 3 == 1 and n > 4:
            Q, _ = np.linalg.qr(X)
            ev = np.repeat(rng.standard_normal((n + 2) // 3), 3)[:n]
            X = (Q * ev) @ Q.T
        elif k % 3 == 2 and n > 4:
            Y = X[:, : n // 4]
            X = Y @ Y.T
        else:
            X = X + X.T
        mats.append(X)
    a_offs = np.cumsum([0] + [n * n for n in ns])[:-1]
    w_offs = np.cumsum([0] + ns)[:-1]
    flat = np.concatenate([m.ravel(order="F") for m in mats] + [np.zeros(1)])
    dA = dev.put(flat)
    dw = dev.empty(sum(ns) + 1)
    ctx.syev_batched(ns, a_offs, dA.data_ptr(), w_offs, dw.data_ptr())
    A = dev.get(dA, (flat.size,))
    w = dev.get(dw, (sum(ns) + 1,))
    for X, n, ao, wo in zip(mats, ns, a_offs, w_offs):
        if n == 0:
            continue
        U = A[ao: ao + n * n].reshape((n, n), order="F")
        wn = w[wo: wo + n]
        wref = np.linalg.eigvalsh(X)[::-1]
        scale = max(1.0, np.max(np.abs(wref)))
        assert np.max(np.abs(wn - wref)) / scale < 1e-12, n
        assert np.max(np.abs(U.T @ U - np.eye(n))) < 1e-12, n
        assert np.max(np.abs(U @ np.diag(wn) @ U.T - X)) / scale < 1e-12, n
