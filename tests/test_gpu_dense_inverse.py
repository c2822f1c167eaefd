"""The coarse-space Gauss-Jordan inverse (blocked, partial pivoting, rank-b
updates on the FP64 matrix cores) against numpy on dense matrices whose
sizes cover every panel shape (one, two and four rows per thread; partial last
panels; sizes below one MFMA tile) and matrices that need row interchanges."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(rt, abi):
    c = rt.Context(abi.PRECISION_FP64)
    yield c
    c.close()


def well_conditioned(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return a + n ** 0.5 * 2.0 * np.eye(n)


@pytest.mark.parametrize("n", [1, 2, 3, 15, 16, 17, 33, 100, 729, 1024, 1025, 1500, 2049, 3000])
def test_dense_inverse_matches_numpy(ctx, n):
    import torch
    a = well_conditioned(n, seed=n)
    x = ctx.dense_inverse(torch.from_numpy(a).cuda()).cpu().numpy()
    ref = np.linalg.inv(a)
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    res = np.linalg.norm(a @ x - np.eye(n)) / np.sqrt(n)
    print(n, err, res, flush=True)
    assert err <= 1e-12 and res <= 1e-12


@pytest.mark.parametrize("n", [8, 64, 729])
def test_dense_inverse_pivots(ctx, n):
    """A row permutation of a diagonally dominant matrix: without the row
    interchanges the first pivots are (near) zero."""
    import torch
    rng = np.random.default_rng(7)
    d = rng.standard_normal((n, n)) * 0.1 + np.diag(rng.uniform(1.0, 2.0, n))
    perm = np.roll(np.arange(n), n // 2 + 1)
    a = d[perm]
    a[0, 0] = 0.0  # a structurally zero first pivot
    x = ctx.dense_inverse(torch.from_numpy(a).cuda()).cpu().numpy()
    ref = np.linalg.inv(a)
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    assert err <= 1e-12, err


def test_dense_inverse_singular_raises(ctx, rt):
    """A zero pivot in the second panel (step 34 of 40): the later steps never
    run, so the pivot order must not read their unwritten entries.  The call
    raises, and the context stays usable."""
    import torch
    a = np.zeros((40, 40))
    a[:33, :33] = 2.0 * np.eye(33)
    a[33:, 33:] = 1.0  # exact arithmetic: step 34 meets an exact zero column
    with pytest.raises(rt.LsmpsError) as e:
        ctx.dense_inverse(torch.from_numpy(a).cuda())
    assert "zero pivot" in str(e.value)
    b = well_conditioned(50, seed=4)
    x = ctx.dense_inverse(torch.from_numpy(b).cuda()).cpu().numpy()
    assert np.linalg.norm(b @ x - np.eye(50)) <= 1e-10
