"""The oracle's PPE solvers against a direct sparse LU, and its Schwarz
preconditioner against dense block solves done independently with LAPACK.

The PPE of a dam-break corner has NearSurface rows with a positive diagonal
(wall-dominated LSMPS Laplacian rows); they seed the Schwarz blocks."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


@pytest.fixture(scope="module")
def system(oracle, cases, abi):
    case = cases.dam_break(block=(24, 12, 12), tank=(48, 12, 24), l0=0.01, seed=0)
    opt = abi.default_solver_options()
    res = oracle.pipeline(case, opt)
    ppe = res["ppe"]
    n = case.n_fluid
    A = sp.csr_matrix((ppe["values"], ppe["col_indices"], ppe["row_offsets"]), shape=(n, n))
    return A, ppe["rhs"].copy(), ppe


@pytest.fixture(scope="module")
def direct(system):
    A, b, _ = system
    return spla.spsolve(A.tocsc(), b)


def _blocks_numpy(A, max_block=256):
    """Independent statement of the block definition: seeds are rows with a
    positive diagonal and off-diagonal couplings; a block is the seed's 2-hop set
    of coupled rows (1-hop if that exceeds max_block)."""
    A = A.tocsr()
    n = A.shape[0]
    d = A.diagonal()
    off = np.asarray(abs(A - sp.diags(d)).sum(axis=1)).ravel()
    eligible = off > 0
    seeds = np.nonzero(eligible & (d > 0))[0]
    blocks = []
    for s in seeds:
        one = np.unique(A.indices[A.indptr[s]:A.indptr[s + 1]])
        one = one[eligible[one]]
        two = np.unique(np.concatenate([one] + [A.indices[A.indptr[j]:A.indptr[j + 1]] for j in one]))
        two = two[eligible[two]]
        blocks.append(two if two.size <= max_block else one[:max_block])
    return blocks, d


def test_positive_diagonal_rows_exist(system):
    A, _, _ = system
    d = A.diagonal()
    off = np.asarray(abs(A - sp.diags(d)).sum(axis=1)).ravel()
    assert np.count_nonzero((d > 0) & (off > 0)) > 0


def test_schwarz_apply_matches_lapack_blocks(oracle, system):
    A, _, ppe = system
    n = A.shape[0]
    v = np.random.default_rng(7).standard_normal(n)
    z, nb, nrows = oracle.schwarz_apply(ppe["row_offsets"], ppe["col_indices"], ppe["values"], v)
    blocks, d = _blocks_numpy(A)
    assert nb == len(blocks) and nb > 0
    assert nrows == sum(b.size for b in blocks)
    ref = v / np.where(d != 0, d, 1.0)
    acc = np.zeros(n)
    cnt = np.zeros(n)
    Ad = A.tocsr()
    for b in blocks:
        acc[b] += np.linalg.solve(Ad[b][:, b].toarray(), v[b])
        cnt[b] += 1
    cov = cnt > 0
    ref[cov] = acc[cov] / cnt[cov]
    assert np.linalg.norm(z - ref) <= 1e-10 * np.linalg.norm(ref)


@pytest.mark.parametrize("method,prec", [("IDRS", "SCHWARZ"), ("IDRS", "JACOBI"),
                                         ("BICGSTAB", "JACOBI")])
def test_solver_matches_direct_lu(oracle, abi, system, direct, method, prec):
    A, b, ppe = system
    x_direct = direct
    opt = abi.default_solver_options()
    opt.method = getattr(abi, "SOLVER_" + method)
    opt.preconditioner = getattr(abi, "PRECOND_" + prec)
    opt.clamp_negative = 0
    x, rep = oracle.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], b, opt)
    assert rep["converged"] == 1
    assert rep["relative_residual"] <= 1e-10
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b) * 1.0001
    assert np.linalg.norm(x - x_direct) <= 1e-7 * np.linalg.norm(x_direct)


def test_schwarz_reduces_iterations(oracle, abi, system):
    _, b, ppe = system
    its = {}
    for prec in ("JACOBI", "SCHWARZ"):
        opt = abi.default_solver_options()
        opt.preconditioner = getattr(abi, "PRECOND_" + prec)
        _, rep = oracle.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], b, opt)
        its[prec] = rep["iterations"]
        assert rep["converged"] == 1
    assert its["SCHWARZ"] < its["JACOBI"], its
