"""GPU FP64 subspace solvers (derek_hip_heev / derek_hip_hegv) against numpy LAPACK."""
import numpy as np
import pytest

from derek_amd import lib

pytestmark = pytest.mark.gpu


def herm(rng, n):
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return 0.5 * (a + a.conj().T)


@pytest.fixture(scope="module")
def ctx():
    c = lib.Context((8, 8, 8), 1, 4, 1.0, 1)
    yield c
    c.close()


# sizes on both sides of the switch from the in-LDS reduction (up to 98
# columns) to multi-workgroup steps before a 98-column tail
@pytest.mark.parametrize("n,nev", [(1, 1), (2, 2), (3, 2), (16, 16), (17, 17), (33, 20), (64, 64), (80, 80), (97, 97),
                                   (98, 98), (99, 99), (128, 128), (160, 80), (161, 161), (200, 100), (300, 150),
                                   (600, 300)])
def test_heev_random(ctx, n, nev):
    print("heev", n, nev, flush=True)
    rng = np.random.default_rng(n)
    A = herm(rng, n)
    w, Z = ctx.heev(A, nev)
    ref = np.linalg.eigvalsh(A)[:nev]
    scale = max(1.0, np.abs(ref).max())
    assert np.max(np.abs(w - ref)) < 1e-12 * scale * n
    res = np.max(np.abs(A @ Z - Z * w[None, :]))
    assert res < 1e-11 * scale * np.sqrt(n), res
    orth = np.max(np.abs(Z.conj().T @ Z - np.eye(nev)))
    assert orth < 1e-11 * np.sqrt(n), orth


def test_heev_degenerate_and_diagonal(ctx):
    rng = np.random.default_rng(5)
    n = 40
    lam = np.repeat([-1.0, 0.5, 2.0, 2.0000000001, 3.0], 8)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    A = (Q * lam) @ Q.conj().T
    w, Z = ctx.heev(A, n)
    assert np.max(np.abs(w - np.sort(lam))) < 1e-12
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(n))) < 1e-10
    assert np.max(np.abs(A @ Z - Z * w[None, :])) < 1e-10
    D = np.diag(np.linspace(3.0, -2.0, 12)).astype(complex)
    w, Z = ctx.heev(D, 5)
    assert np.allclose(w, np.sort(np.diag(D).real)[:5], atol=1e-14)
    assert np.max(np.abs(D @ Z - Z * w[None, :])) < 1e-12


@pytest.mark.parametrize("n", [60, 150, 400])
def test_heev_columns_without_reflector(ctx, n):
    """Columns already zero below the subdiagonal take no reflector, both in
    the in-LDS reduction (n <= 98) and in the multi-workgroup one."""
    print("heev skip", n, flush=True)
    rng = np.random.default_rng(11 + n)
    A = herm(rng, n)
    # a complex tridiagonal block followed by a dense block: columns 0..n/2-2
    # have nothing below their subdiagonal
    h = n // 2
    A[:h, :h] = 0
    A[:h, h:] = 0
    A[h:, :h] = 0
    idx = np.arange(h)
    A[idx, idx] = rng.normal(size=h)
    off = rng.normal(size=h - 1) + 1j * rng.normal(size=h - 1)
    A[idx[1:], idx[:-1]] = off
    A[idx[:-1], idx[1:]] = off.conj()
    w, Z = ctx.heev(A, n)
    ref = np.linalg.eigvalsh(A)
    assert np.max(np.abs(w - ref)) < 1e-12 * n * max(1.0, np.abs(ref).max())
    assert np.max(np.abs(A @ Z - Z * w[None, :])) < 1e-11 * np.sqrt(n) * max(1.0, np.abs(ref).max())
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(n))) < 1e-11 * np.sqrt(n)


@pytest.mark.parametrize("m,nev", [(20, 10), (160, 80)])
def test_hegv(ctx, m, nev):
    rng = np.random.default_rng(m)
    B = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    S = np.eye(m) + 0.05 * (B.conj().T @ B) / m
    H = herm(rng, m)
    w, Cm = ctx.hegv(S, H, nev)
    L = np.linalg.cholesky(S)
    Li = np.linalg.inv(L)
    ref = np.linalg.eigvalsh(Li @ H @ Li.conj().T)[:nev]
    assert np.max(np.abs(w - ref)) < 1e-11 * max(1, np.abs(ref).max())
    assert np.max(np.abs(Cm.conj().T @ S @ Cm - np.eye(nev))) < 1e-10
    assert np.max(np.abs(H @ Cm - S @ Cm * w[None, :])) < 1e-10


def test_hegv_drops_dependent_columns(ctx):
    rng = np.random.default_rng(9)
    npw, m = 60, 12
    Y = rng.normal(size=(npw, m)) + 1j * rng.normal(size=(npw, m))
    Y[:, 7] = Y[:, 2] * (1 + 1e-13)  # numerically dependent on an earlier column
    Hop = herm(rng, npw)
    S = Y.conj().T @ Y
    H = Y.conj().T @ Hop @ Y
    w, Cm = ctx.hegv(S, H, 5)
    assert np.all(Cm[7] == 0)
    keep = [i for i in range(m) if i != 7]
    Sa, Ha = S[np.ix_(keep, keep)], H[np.ix_(keep, keep)]
    La = np.linalg.inv(np.linalg.cholesky(Sa))
    ref = np.linalg.eigvalsh(La @ Ha @ La.conj().T)[:5]
    assert np.max(np.abs(w - ref)) < 1e-9 * np.abs(ref).max()
