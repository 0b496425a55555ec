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


@pytest.mark.parametrize("n", [1, 2, 3, 10, 64, 65, 100, 128, 129, 256, 257, 400, 512, 513, 1024, 2048])
def test_syev(ctx, dev, n):
    """Single matrices of every size class (our kernels up to 256 rows:
    reduction LDS-resident up to 64 and up to 128, in global memory above;
    rocSOLVER's dsyevd on the negated, padded block above): eigenvalues vs
    numpy, orthonormal vectors, reconstruction, descending order."""
    rng = np.random.default_rng(1000 + n)
    X = rng.standard_normal((n, n))
    if n in (64, 128, 256, 512, 2048):
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
    ns = [5, 64, 1, 90, 200, 33, 128, 0, 150, 300, 256, 2, 520]
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


def test_syev_batched_host_eigenvalues(ctx, dev):
    """itx_syev_batched_host hands the host the same eigenvalues (bit for
    bit) that it writes to w, with the same eigenvectors as the plain
    batched call."""
    rng = np.random.default_rng(78)
    ns = [40, 0, 100, 130, 260, 3]
    mats = []
    for n in ns:
        X = rng.standard_normal((n, max(n // 3, 1)))
        mats.append(X @ X.T)
    a_offs = np.cumsum([0] + [n * n for n in ns])[:-1]
    w_offs = np.cumsum([0] + ns)[:-1]
    flat = np.concatenate([m.ravel(order="F") for m in mats] + [np.zeros(1)])
    dA, dA2 = dev.put(flat), dev.put(flat)
    dw, dw2 = dev.empty(sum(ns) + 1), dev.empty(sum(ns) + 1)
    wh = ctx.syev_batched_host(ns, a_offs, dA.data_ptr(), w_offs, dw.data_ptr())
    ctx.syev_batched(ns, a_offs, dA2.data_ptr(), w_offs, dw2.data_ptr())
    w, w2 = dev.get(dw, (sum(ns) + 1,)), dev.get(dw2, (sum(ns) + 1,))
    assert np.array_equal(wh[: sum(ns)], w[: sum(ns)])
    assert np.array_equal(w[: sum(ns)], w2[: sum(ns)])
    assert np.array_equal(dev.get(dA, (flat.size,)), dev.get(dA2, (flat.size,)))


@pytest.mark.parametrize("m,n,kind", [(1, 1, "rand"), (1, 7, "rand"), (7, 1, "rand"), (5, 9, "rand"),
                                      (9, 5, "rand"), (40, 40, "rank"), (64, 130, "rand"), (130, 64, "rand"),
                                      (100, 100, "decay"), (33, 200, "decay")])
def test_gesvd(ctx, dev, m, n, kind):
    """One-sided Jacobi SVD vs numpy: singular values (descending) to 1e-12
    of the largest, orthonormal singular vectors, reconstruction to 1e-13
    relative -- the accuracy stress test of matrix_test.cc:1808-1841 is
    the 'decay' case (spectrum 0.8^j)."""
    rng = np.random.default_rng(m * 1000 + n)
    M = rng.standard_normal((m, n))
    k = min(m, n)
    if kind == "rank":
        M = rng.standard_normal((m, 7)) @ rng.standard_normal((7, n))
    elif kind == "decay":
        Uq, _ = np.linalg.qr(rng.standard_normal((m, k)))
        Vq, _ = np.linalg.qr(rng.standard_normal((n, k)))
        M = (Uq * 0.8 ** np.arange(k)) @ Vq.T
    dA = dev.put(M)
    dU, dS, dV = dev.empty(m * k), dev.empty(k), dev.empty(n * k)
    ctx.gesvd(m, n, dA.data_ptr(), dU.data_ptr(), dS.data_ptr(), dV.data_ptr())
    U, S, V = dev.get(dU, (m, k)), dev.get(dS, (k,)), dev.get(dV, (n, k))
    ref = np.linalg.svd(M, compute_uv=False)
    assert np.all(np.diff(S) <= 0)
    assert np.max(np.abs(S - ref)) < 1e-12 * max(1.0, ref[0])
    assert np.linalg.norm(U @ np.diag(S) @ V.T - M) / np.linalg.norm(M) < 1e-13
    nz = S > 1e-10 * S[0]
    for X in (U[:, nz], V[:, nz]):
        assert np.max(np.abs(X.T @ X - np.eye(X.shape[1]))) < 1e-12
    assert np.array_equal(dev.get(dA, (m, n)), M)  # input untouched


def test_gesvd_batched(ctx, dev):
    """Several blocks of different shapes in one launch, stored back to back."""
    rng = np.random.default_rng(5)
    shapes = [(3, 5), (8, 2), (1, 1), (20, 20), (0, 4), (6, 6)]
    mats = [rng.standard_normal(s) for s in shapes]
    ks = [min(s) for s in shapes]
    a_offs = np.cumsum([0] + [a * b for a, b in shapes])[:-1]
    u_offs = np.cumsum([0] + [a * k for (a, _), k in zip(shapes, ks)])[:-1]
    v_offs = np.cumsum([0] + [b * k for (_, b), k in zip(shapes, ks)])[:-1]
    s_offs = np.cumsum([0] + ks)[:-1]
    A = np.concatenate([x.ravel(order="F") for x in mats] + [np.zeros(1)])
    dA = dev.put(A)
    nU = sum(a * k for (a, _), k in zip(shapes, ks)) + 1
    nV = sum(b * k for (_, b), k in zip(shapes, ks)) + 1
    dU, dS, dV = dev.empty(nU), dev.empty(sum(ks) + 1), dev.empty(nV)
    ctx.gesvd_batched([s[0] for s in shapes], [s[1] for s in shapes], a_offs, dA.data_ptr(), u_offs,
                      dU.data_ptr(), s_offs, dS.data_ptr(), v_offs, dV.data_ptr())
    Uh, Sh, Vh = dev.get(dU, (nU,)), dev.get(dS, (sum(ks) + 1,)), dev.get(dV, (nV,))
    for M, (a, b), k, uo, so, vo in zip(mats, shapes, ks, u_offs, s_offs, v_offs):
        if k == 0:
            continue
        U = Uh[uo: uo + a * k].reshape((a, k), order="F")
        V = Vh[vo: vo + b * k].reshape((b, k), order="F")
        S = Sh[so: so + k]
        assert np.linalg.norm(U @ np.diag(S) @ V.T - M) < 1e-12 * max(1.0, np.linalg.norm(M))


def test_syev_batched_large_size_classes(ctx, dev):
    """Several blocks above 256 rows in one call: they go to the large-block
    path in size classes (padded slots of 320, 512, 704 and 1024 rows here),
    beside small blocks on our kernels; density-matrix-like blocks with
    exact zero eigenvalues included.  Every block against numpy."""
    rng = np.random.default_rng(79)
    ns = [300, 20, 301, 257, 700, 1024, 0, 500, 450, 130]
    mats = []
    for k, n in enumerate(ns):
        if k % 2:
            Y = rng.standard_normal((n, max(1, n // 3)))
            X = Y @ Y.T
        else:
            X = rng.standard_normal((n, n))
            X = X + X.T
        mats.append(X)
    a_offs = np.cumsum([0] + [n * n for n in ns])[:-1]
    w_offs = np.cumsum([0] + ns)[:-1]
    flat = np.concatenate([m.ravel(order="F") for m in mats] + [np.zeros(1)])
    dA = dev.put(flat)
    dw = dev.empty(sum(ns) + 1)
    wh = ctx.syev_batched_host(ns, a_offs, dA.data_ptr(), w_offs, dw.data_ptr())
    A = dev.get(dA, (flat.size,))
    w = dev.get(dw, (sum(ns) + 1,))
    assert np.array_equal(wh[: sum(ns)], w[: sum(ns)])
    for X, n, ao, wo in zip(mats, ns, a_offs, w_offs):
        if n == 0:
            continue
        U = A[ao: ao + n * n].reshape((n, n), order="F")
        wn = w[wo: wo + n]
        wref = np.linalg.eigvalsh(X)[::-1]
        scale = max(1.0, np.max(np.abs(wref)))
        assert np.max(np.abs(wn - wref)) / scale < 1e-12, n
        assert np.all(np.diff(wn) <= 0), n
        assert np.max(np.abs(U.T @ U - np.eye(n))) < 1e-12, n
        assert np.max(np.abs(U @ np.diag(wn) @ U.T - X)) / scale < 1e-12, n


def test_syev_rejects_blocks_above_the_limit(ctx, dev):
    """Blocks above 8192 rows: ITX_ERR_UNSUPPORTED before any device work
    (the buffers are never touched)."""
    from itensor_amd.itx import ItxError

    dA, dw = dev.empty(4), dev.empty(4)
    with pytest.raises(ItxError, match=r"\(4\).*8192"):
        ctx.syev(8193, dA.data_ptr(), dw.data_ptr())
