"""The PPE solve's SpMV layout (diagonals of 256-row groups read as x windows,
the other entries in sliced ELLPACK) against a CSR product in numpy, through
lsmps_ppe_spmv: a real PPE matrix of a jittered dam-break, random unstructured
matrices (no diagonals at all), banded matrices with runs of consecutive
diagonals longer than one window holds and offsets that reach outside [0, n)
near both ends (the value-by-value class), and sizes around the 64-row chunk
and 256-row group boundaries.  FP64 values must agree to rounding of the sum;
with FP32 values the product must equal the CSR product of the FP32-rounded
values, accumulated in FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(rt, abi):
    c = rt.Context(abi.PRECISION_FP64)
    yield c
    c.close()


def csr_product(rows, cols, vals, x):
    n = rows.shape[0] - 1
    y = np.zeros(n)
    mag = np.zeros(n)
    r = np.repeat(np.arange(n), np.diff(rows))
    np.add.at(y, r, vals * x[cols])
    np.add.at(mag, r, np.abs(vals * x[cols]))
    return y, mag


def check(ctx, abi, rows, cols, vals, seed, label):
    import torch
    n = rows.shape[0] - 1
    x = np.random.default_rng(seed).standard_normal(n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    for prec in (abi.PRECISION_FP64, abi.PRECISION_MIXED):
        v = vals if prec == abi.PRECISION_FP64 else vals.astype(np.float32).astype(np.float64)
        ref, mag = csr_product(rows, cols, v, x)
        y = ctx.ppe_spmv(dev(rows), dev(cols), dev(vals), dev(x), prec).cpu().numpy()
        bad = np.abs(y - ref) > 1e-13 * mag + 1e-300
        print(label, "precision", prec, "n", n, "nnz", vals.shape[0],
              "max err / magnitude", float(np.max(np.abs(y - ref) / np.maximum(mag, 1e-300))),
              flush=True)
        assert not bad.any(), (label, prec, np.flatnonzero(bad)[:10])


def random_csr(n, per_row, band, seed, diagonal=True):
    """Rows of distinct columns; with `band` > 0 the columns lie within
    +-band of the row (clipped to [0, n)), else anywhere."""
    rng = np.random.default_rng(seed)
    rows = [0]
    cols = []
    for i in range(n):
        k = int(rng.integers(0, per_row + 1))
        if band > 0:
            lo, hi = max(0, i - band), min(n, i + band + 1)
            c = rng.choice(np.arange(lo, hi), size=min(k, hi - lo), replace=False)
        else:
            c = rng.choice(n, size=min(k, n), replace=False)
        c = set(int(v) for v in c)
        if diagonal:
            c.add(i)
        c = sorted(c)
        cols.extend(c)
        rows.append(len(cols))
    cols = np.array(cols, np.int32)
    vals = rng.standard_normal(cols.shape[0])
    return np.array(rows, np.int32), cols, vals


def stencil_csr(n, offsets, seed, drop=0.0):
    """Every row takes the columns i + d for d in offsets inside [0, n); a
    fraction `drop` of the entries is removed at random."""
    rng = np.random.default_rng(seed)
    offsets = np.array(sorted(offsets))
    i = np.repeat(np.arange(n), offsets.shape[0])
    j = i + np.tile(offsets, n)
    keep = (j >= 0) & (j < n) & (rng.random(j.shape[0]) >= drop)
    i, j = i[keep], j[keep]
    rows = np.zeros(n + 1, np.int64)
    np.add.at(rows, i + 1, 1)
    rows = np.cumsum(rows).astype(np.int32)
    vals = rng.standard_normal(j.shape[0])
    return rows, j.astype(np.int32), vals


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 256, 257, 300, 1000, 4097])
def test_spmv_small_sizes(ctx, abi, n):
    rows, cols, vals = random_csr(n, 12, 20, seed=n)
    check(ctx, abi, rows, cols, vals, seed=n, label=f"banded n={n}")


def test_spmv_unstructured(ctx, abi):
    rows, cols, vals = random_csr(20000, 30, 0, seed=1)
    check(ctx, abi, rows, cols, vals, seed=1, label="unstructured")


def test_spmv_long_runs_and_far_offsets(ctx, abi):
    """Runs of 20 and 11 consecutive diagonals (cut at 8 per window), isolated
    diagonals, and offsets of +-9000 on 20,000 rows, whose windows leave [0, n)
    in the first and last groups they touch."""
    offs = list(range(-10, 10)) + list(range(300, 311)) + [-301, 57, -9000, 9000, 9001]
    rows, cols, vals = stencil_csr(20000, offs, seed=2, drop=0.05)
    check(ctx, abi, rows, cols, vals, seed=2, label="long runs")


def test_spmv_lattice_ppe(ctx, abi, cases, oracle):
    """The assembled PPE of a jittered 24x24x24 dam-break block (lattice order:
    ~123 shared offsets per group, wall-adjacent rows, Dirichlet rows)."""
    case = cases.dam_break(block=(24, 24, 24), tank=(48, 24, 48), l0=0.005, seed=3)
    p = oracle.pipeline(case)["ppe"]
    rows = np.asarray(p["row_offsets"], np.int32)
    n = rows.shape[0] - 1
    used = int(rows[n])
    cols = np.asarray(p["col_indices"][:used], np.int32)
    vals = np.asarray(p["values"][:used], np.float64)
    check(ctx, abi, rows, cols, vals, seed=3, label="lattice PPE")
