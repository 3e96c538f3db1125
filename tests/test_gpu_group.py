"""Grouped LU of uneven-order problems in lockstep (mg_group_*) against
LAPACK's zgetrf / cgetrf (scipy's lu_factor: the same pivoting rule, the
largest |re| + |im|, first on ties), against the per-problem pivoted LU of
the library, and against solves in complex128."""
import numpy as np
import pytest
import scipy.linalg as sl
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import engine as E

DEV = "cuda:0"
# orders around the leaf's on-chip rows (576 in LDS, then register slots, then
# streamed rows: 1088 complex128 / 2176 complex64 rows at 1024 threads)
SIZES = [700, 33, 1, 256, 1029, 130, 64, 65, 17, 128, 129, 300, 1100, 2300]


def say(*a):
    print(*a, flush=True)


def _mats(sizes, seed, dtype=np.complex128, shrink=0.05):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        A[np.arange(n), np.arange(n)] *= shrink     # small diagonal: every column pivots
        out.append(A.astype(dtype))
    return out


def _unpack(LU):
    n = LU.shape[0]
    return np.tril(LU, -1) + np.eye(n), np.triu(LU)


@pytest.mark.parametrize("cfg", [(128, 16, 0, 1024), (64, 8, 1, 512), (256, 16, 0, 256), (32, 4, 1, 256)])
def test_group_factors_match_lapack(cfg):
    """Pivots equal LAPACK's for every problem; P A = L U to rounding; the
    factors agree with zgetrf's."""
    nb, leaf, tile, lt = cfg
    say("group getrf vs lapack", cfg)
    host = _mats(SIZES, 1)
    mats = [torch.from_numpy(A).to(DEV) for A in host]
    g = E.GroupLU(mats, prec="c128", nb=nb, leaf=leaf, tile=tile, leaf_threads=lt).factor()
    torch.cuda.synchronize()
    assert (g.info.cpu() == 0).all()
    for i, A in enumerate(host):
        lu, piv = sl.lu_factor(A)
        mine = mats[i].cpu().numpy()
        ip = g.ipiv(i).cpu().numpy()
        assert np.array_equal(ip, piv), (i, A.shape)
        err = np.abs(mine - lu).max() / np.abs(lu).max()
        # both factorisations round differently (blocking, FMA): ~eps n (2e-12 at n = 2300)
        assert err < 5e-15 * max(A.shape[0], 100), (i, A.shape, err)
        Lm, Um = _unpack(mine)
        PA = A.copy()
        for r, p in enumerate(ip):
            PA[[r, p]] = PA[[p, r]]
        assert np.abs(Lm @ Um - PA).max() <= 1e-12 * np.abs(A).max() * A.shape[0]
    say("launches", g.launches)
    g.close()


def test_group_solve_matches_sequential_and_numpy():
    """Grouped solve of 32 right-hand sides equals the library's per-problem
    pivoted LU solve and numpy's solve."""
    say("group getrs")
    host = _mats(SIZES, 2, shrink=1.0)
    rng = np.random.default_rng(3)
    rhs_host = [rng.normal(size=(A.shape[0], 32)) + 1j * rng.normal(size=(A.shape[0], 32)) for A in host]
    mats = [torch.from_numpy(A).to(DEV) for A in host]
    rhs = [torch.from_numpy(B).to(DEV) for B in rhs_host]
    g = E.GroupLU(mats, rhs, prec="c128")
    g.factor()
    g.solve()
    torch.cuda.synchronize()
    for i, (A, B) in enumerate(zip(host, rhs_host)):
        X = rhs[i].cpu().numpy()
        ref = np.linalg.solve(A, B)
        err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
        assert err < 1e-11, (i, err)
        F = E.Factor(torch.from_numpy(A).to(DEV))
        Y = torch.from_numpy(B).to(DEV)
        F.solve(Y)
        seq = Y.cpu().numpy()
        # two backward-stable solves differ by ~eps cond(A): up to ~1e3 here
        assert np.linalg.norm(X - seq) / np.linalg.norm(seq) < 1e-11, i
    g.close()


def test_group_zero_pivot_info():
    """A zero column makes info the first zero pivot (1-based), as zgetrf,
    and leaves the other problems' factors untouched by it."""
    say("group info")
    host = _mats([90, 200, 40], 4, shrink=1.0)
    host[1][:, 77] = 0
    mats = [torch.from_numpy(A).to(DEV) for A in host]
    g = E.GroupLU(mats, prec="c128").factor()
    torch.cuda.synchronize()
    info = g.info.cpu().numpy()
    assert info[0] == 0 and info[2] == 0
    assert info[1] == 78
    lu, piv = sl.lu_factor(host[0])
    assert np.array_equal(g.ipiv(0).cpu().numpy(), piv)
    g.close()


@pytest.mark.parametrize("prec,tol,nb,pivot,nrhs", [("c64", 2e-4, 256, "partial", 8),
                                                    ("bf16x3", 2e-3, 256, "tournament", 8),
                                                    ("bf16x3", 2e-3, 128, "tournament", 8),
                                                    ("bf16x3", 2e-3, 384, "tournament", 40),
                                                    ("bf16x3", 2e-3, 64, "tournament", 8),
                                                    ("bf16x3", 2e-3, 256, "partial", 40)])
def test_group_complex_float(prec, tol, nb, pivot, nrhs):
    """complex64 factors (the trailing updates on FP32 or BF16x3 products;
    BF16x3 at panel widths 128, 256 and 384 from operands split once per
    step, at 64 split in every tile; tournament or partial pivoting in the
    leaves; BF16x3 solves from inverted diagonal blocks, one and two
    32-column tiles): solutions within the precision of the factor of the
    complex128 ones."""
    say("group", prec, nb, pivot, nrhs)
    host = _mats(SIZES, 5, dtype=np.complex64, shrink=1.0)
    rng = np.random.default_rng(6)
    rhs_host = [(rng.normal(size=(A.shape[0], nrhs)) + 1j * rng.normal(size=(A.shape[0], nrhs))).astype(np.complex64)
                for A in host]
    for i in range(len(host)):
        host[i] = (host[i] + 4 * np.sqrt(host[i].shape[0]) * np.eye(host[i].shape[0])).astype(np.complex64)
    mats = [torch.from_numpy(A).to(DEV) for A in host]
    rhs = [torch.from_numpy(B).to(DEV) for B in rhs_host]
    g = E.GroupLU(mats, rhs, prec=prec, nb=nb, pivot=pivot)
    g.factor()
    g.solve()
    torch.cuda.synchronize()
    assert (g.info.cpu() == 0).all()
    for i, (A, B) in enumerate(zip(host, rhs_host)):
        ref = np.linalg.solve(A.astype(np.complex128), B.astype(np.complex128))
        X = rhs[i].cpu().numpy()
        err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
        assert err < tol, (i, err)
    g.close()


@pytest.mark.parametrize("nrhs", [5, 40])
def test_group_residual_and_colnorms(nrhs):
    """B_i -= A_i X_i over the group in one launch and the column norms,
    against numpy, complex128 (orders off the 16-deep k chunks and 128-row
    blocks, one and two 32-column tiles) and complex64."""
    say("group residual", nrhs)
    for dt, tol in ((np.complex128, 1e-13), (np.complex64, 1e-5)):
        host = _mats([300, 41, 1, 129, 1040], 7, dtype=dt, shrink=1.0)
        rng = np.random.default_rng(8)
        Bh = [(rng.normal(size=(A.shape[0], nrhs)) + 1j * rng.normal(size=(A.shape[0], nrhs))).astype(dt)
              for A in host]
        Xh = [(rng.normal(size=(A.shape[0], nrhs)) + 1j * rng.normal(size=(A.shape[0], nrhs))).astype(dt)
              for A in host]
        mats = [torch.from_numpy(A).to(DEV) for A in host]
        rhs = [torch.from_numpy(B).to(DEV) for B in Bh]
        g = E.GroupLU(mats, rhs, prec="c128" if dt == np.complex128 else "c64")
        g.set_aux([torch.from_numpy(X).to(DEV) for X in Xh])
        g.residual()
        nrm = g.colnorms().cpu().numpy()
        for i, (A, B, X) in enumerate(zip(host, Bh, Xh)):
            ref = B.astype(np.complex128) - A.astype(np.complex128) @ X.astype(np.complex128)
            got = rhs[i].cpu().numpy()
            assert np.abs(got - ref).max() <= tol * np.abs(A).max() * A.shape[0], i
            assert np.allclose(nrm[i], np.linalg.norm(got.astype(np.complex128), axis=0), rtol=1e-12), i
        g.close()


def _gepp_labels(M, labels, w):
    """Partial pivoting (largest |re| + |im|, first on ties) on the first w
    columns of M (complex128): the labels of the pivot rows, in order."""
    M = M[:, :w].astype(np.complex128).copy()
    labels = list(labels)
    for j in range(min(w, M.shape[0])):
        a = np.abs(M[j:, j].real) + np.abs(M[j:, j].imag)
        p = j + int(np.argmax(a))
        M[[j, p]] = M[[p, j]]
        labels[j], labels[p] = labels[p], labels[j]
        if M[j, j] != 0:
            M[j + 1:, j] /= M[j, j]
            M[j + 1:, j + 1:] -= np.outer(M[j + 1:, j], M[j, j + 1:])
    return labels[:w]


def _tournament_first_leaf(A, w=16, rows=256):
    """The tournament rule on columns [0, w): every chunk of `rows` rows
    offers the rows partial pivoting picks in it, then partial pivoting among
    the offered rows (original values) picks the leaf's pivot rows in order;
    returned as getrf's swap sequence."""
    n = A.shape[0]
    cand = []
    for r0 in range(0, n, rows):
        cand += _gepp_labels(A[r0:r0 + rows], range(r0, min(n, r0 + rows)), w)
    sel = _gepp_labels(A[cand], cand, w)
    cur = list(range(n))
    ipiv = []
    for j, row in enumerate(sel):
        q = cur.index(row)
        ipiv.append(q)
        cur[j], cur[q] = cur[q], cur[j]
    return np.array(ipiv)


def test_group_tournament_leaf():
    """BF16x3 tournament leaves: the first leaf's pivots equal the rule's
    (restated in numpy on the same complex64 data), every pivot is a valid
    swap, P A = L U to the factor's precision, and a zero column gives
    getrf's info."""
    say("group tournament")
    host = _mats([2300, 700, 1029, 300, 17], 9, dtype=np.complex64, shrink=0.05)
    errs = {}
    for pivot in ("partial", "tournament"):
        mats = [torch.from_numpy(A).to(DEV) for A in host]
        g = E.GroupLU(mats, prec="bf16x3", pivot=pivot).factor()
        torch.cuda.synchronize()
        assert (g.info.cpu() == 0).all()
        for i, A in enumerate(host):
            n = A.shape[0]
            ip = g.ipiv(i).cpu().numpy()
            assert ((ip >= np.arange(n)) & (ip < n)).all(), i
            if pivot == "tournament":
                w = min(16, n)
                assert np.array_equal(ip[:w], _tournament_first_leaf(A, w)[:w]), i
            Lm, Um = _unpack(mats[i].cpu().numpy().astype(np.complex128))
            PA = A.astype(np.complex128).copy()
            for r, p in enumerate(ip):
                PA[[r, p]] = PA[[p, r]]
            errs[pivot, i] = np.abs(Lm @ Um - PA).max() / np.abs(A).max()
            say(pivot, i, n, "|L| max %.2f" % np.abs(Lm).max(), "P A - L U %.2e" % errs[pivot, i])
        g.close()
    # BF16x3 products: ~2e-5 per product, ~sqrt(n) of them per entry; the
    # tournament's factors as accurate as partial pivoting's
    for i in range(len(host)):
        assert errs["tournament", i] < 5e-3 and errs["tournament", i] <= 3 * errs["partial", i] + 1e-5, i
    host = _mats([90, 200, 40], 4, dtype=np.complex64, shrink=1.0)
    host[1][:, 77] = 0
    mats = [torch.from_numpy(A).to(DEV) for A in host]
    g = E.GroupLU(mats, prec="bf16x3", pivot="tournament").factor()
    torch.cuda.synchronize()
    info = g.info.cpu().numpy()
    assert info[0] == 0 and info[2] == 0 and info[1] == 78, info
    g.close()
