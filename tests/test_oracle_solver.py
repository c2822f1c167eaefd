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


@pytest.fixture(scope="module")
def coarse_system(oracle, cases, abi):
    """A larger dam-break PPE (23,040 rows) with its fluid cells and labels, so
    that 2x2x2-cell aggregates give a coarse space of about a hundred unknowns
    and the smooth error components it removes matter to the iteration count."""
    case = cases.dam_break(block=(40, 24, 24), tank=(80, 24, 48), l0=0.01, seed=2)
    res = oracle.pipeline(case, abi.default_solver_options())
    ppe = res["ppe"]
    n = case.n_fluid
    A = sp.csr_matrix((ppe["values"], ppe["col_indices"], ppe["row_offsets"]), shape=(n, n))
    agg, count = oracle.coarse_aggregation(case, res["neighbors"], res["surface"]["surface_type"],
                                           factor=2)
    return case, A, ppe, agg, count, res["surface"]["surface_type"]


def test_coarse_aggregation_definition(coarse_system):
    """Aggregates: Inner / NearSurface rows only, grouped by the 2x2x2 block of
    their cell; every aggregate is non-empty and numbered by block order."""
    case, _, _, agg, count, labels = coarse_system
    active = (labels == 0) | (labels == 1)
    assert np.all(agg[~active] == -1) and np.all(agg[active] >= 0)
    assert count > 8 and np.array_equal(np.unique(agg[active]), np.arange(count))
    cfg = case.config
    h = cfg.cell_size
    org = np.array([cfg.cell_origin[k] for k in range(3)])
    xyz = np.stack([case.fluid[k] for k in ("x", "y", "z")], axis=1)
    blk = np.floor((xyz - org) / h).astype(np.int64) // 2
    for g in range(count):  # same block <=> same aggregate
        members = np.nonzero(agg == g)[0]
        assert np.all(blk[members] == blk[members[0]])
    firsts = [blk[np.nonzero(agg == g)[0][0]] for g in range(count)]
    keys = [(b[2], b[1], b[0]) for b in firsts]
    assert keys == sorted(keys)


def test_coarse_apply_matches_dense_galerkin(oracle, coarse_system):
    """z = P (P^T A P)^-1 P^T v against scipy's Galerkin product and LAPACK."""
    _, A, ppe, agg, count, _ = coarse_system
    n = A.shape[0]
    rows = np.nonzero(agg >= 0)[0]
    P = sp.csr_matrix((np.ones(rows.shape[0]), (rows, agg[rows])), shape=(n, count))
    Ac = (P.T @ A @ P).toarray()
    v = np.random.default_rng(4).standard_normal(n)
    ref = P @ np.linalg.solve(Ac, P.T @ v)
    z = oracle.coarse_apply(ppe["row_offsets"], ppe["col_indices"], ppe["values"], v, agg, count)
    assert np.linalg.norm(z - ref) <= 1e-10 * np.linalg.norm(ref)


def test_coarse_space_reduces_iterations(oracle, abi, coarse_system):
    """SCHWARZ_COARSE converges to the direct solution in fewer iterations than
    SCHWARZ on the same system."""
    _, A, ppe, agg, count, _ = coarse_system
    b = ppe["rhs"]
    x_direct = spla.spsolve(A.tocsc(), b)
    its = {}
    for prec in ("SCHWARZ", "SCHWARZ_COARSE"):
        opt = abi.default_solver_options()
        opt.preconditioner = getattr(abi, "PRECOND_" + prec)
        opt.clamp_negative = 0
        x, rep = oracle.solve(ppe["row_offsets"], ppe["col_indices"], ppe["values"], b, opt,
                              coarse=(agg, count))
        assert rep["converged"] == 1
        assert np.linalg.norm(x - x_direct) <= 1e-7 * np.linalg.norm(x_direct)
        assert rep["coarse_unknowns"] == (count if prec == "SCHWARZ_COARSE" else 0)
        its[prec] = rep["iterations"]
    assert its["SCHWARZ_COARSE"] < its["SCHWARZ"], its
