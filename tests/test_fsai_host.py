"""Host construction of the factorized sparse approximate inverse
(nsfem_amd/csrc/host/fsai.cpp) against its definition computed densely per
row with numpy: for row i with lower-colour neighbours L_i,
P_i = L_i + {i}, M[P_i, P_i] g = e_i, g / sqrt(g_i); the factor's values in
the matrix's split layout (G(i, j) on the lower part of row i, G(j, i) on
its higher part)."""

import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fsai") / "fsai_driver")
    host = os.path.join(ROOT, "nsfem_amd", "csrc", "host")
    subprocess.run([gxx, "-O2", "-std=c++20", "-I" + host, "-o", out, os.path.join(HERE, "fsai_driver.cpp"),
                    os.path.join(host, "fsai.cpp"), "-lpthread"], check=True)
    return out


def _split(Mp):
    """Colour-major CSR split into lower part, higher part and diagonal."""
    n = Mp.shape[0]
    rs, spl, col, val = [0], [], [], []
    for i in range(n):
        cs = Mp.indices[Mp.indptr[i]:Mp.indptr[i + 1]]
        vs = Mp.data[Mp.indptr[i]:Mp.indptr[i + 1]]
        for c, v in zip(cs, vs):
            if c < i:
                col.append(c)
                val.append(v)
        spl.append(len(col))
        for c, v in zip(cs, vs):
            if c > i:
                col.append(c)
                val.append(v)
        rs.append(len(col))
    return rs, spl, col, val, Mp.diagonal()


def _run(driver, tmp_path, n, rs, spl, col, val, dg):
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([n, len(col)], np.int32).tofile(f)
        np.array(rs, np.int32).tofile(f)
        np.array(spl, np.int32).tofile(f)
        np.array(col, np.int32).tofile(f)
        np.array(val, np.float64).tofile(f)
        np.asarray(dg, np.float64).tofile(f)
    rc = subprocess.run([driver, fin, fout], capture_output=True, text=True)
    if rc.returncode != 0:
        return rc.returncode, rc.stderr
    out = np.fromfile(fout, np.float64)
    return out[:len(col)], out[len(col):]


def test_fsai_factor_matches_its_definition(driver, tmp_path):
    rng = np.random.default_rng(5)
    n = 600
    # a random symmetric pattern with a dominant diagonal (SPD), put in a
    # greedy colour-major order
    A = sp.random(n, n, density=8 / n, random_state=6, format="csr")
    A = abs(A + A.T)
    A = A + sp.diags(np.asarray(A.sum(axis=1)).ravel() + rng.uniform(0.5, 1.5, n))
    A = sp.csr_matrix(A)
    color = -np.ones(n, int)
    for i in range(n):
        used = set(color[A.indices[A.indptr[i]:A.indptr[i + 1]]])
        c = 0
        while c in used:
            c += 1
        color[i] = c
    perm = np.argsort(color, kind="stable")
    Mp = sp.csr_matrix(A[perm][:, perm])
    Mp.sort_indices()
    rs, spl, col, val, dg = _split(Mp)
    g, gd = _run(driver, tmp_path, n, rs, spl, col, val, dg)
    G = np.zeros((n, n))
    for i in range(n):
        P = list(col[rs[i]:spl[i]]) + [i]
        e = np.zeros(len(P))
        e[-1] = 1.0
        y = np.linalg.solve(Mp[P][:, P].toarray(), e)
        y /= np.sqrt(y[-1])
        G[i, P] = y
    for i in range(n):
        assert abs(gd[i] - G[i, i]) <= 1e-13 * abs(G[i, i])
        for j in range(rs[i], spl[i]):
            assert abs(g[j] - G[i, col[j]]) <= 1e-13 * abs(G[i, i])
        for j in range(spl[i], rs[i + 1]):
            assert abs(g[j] - G[col[j], i]) <= 1e-13 * abs(G[col[j], col[j]])
    # (G M G^T)_ii = 1 and G^T G is a good inverse: its preconditioned
    # matrix's spectrum lies close around 1
    GM = G @ Mp.toarray() @ G.T
    assert np.allclose(np.diag(GM), 1.0, atol=1e-12)
    ev = np.linalg.eigvalsh(GM)
    assert ev[0] > 0.2 and ev[-1] < 2.5


def test_fsai_rejects_an_indefinite_matrix(driver, tmp_path):
    # [[1, 2], [2, 1]] is indefinite: the second row's local system fails
    # (row 0: no lower part, column 1 above; row 1: column 0 below)
    rs, spl, col, val, dg = [0, 1, 2], [0, 2], [1, 0], [2.0, 2.0], np.array([1.0, 1.0])
    rc, err = _run(driver, tmp_path, 2, rs, spl, col, val, dg)
    assert rc == 3 and "not positive definite" in err


def _greedy_colours(A):
    n = A.shape[0]
    color = -np.ones(n, int)
    for i in range(n):
        used = set(color[A.indices[A.indptr[i]:A.indptr[i + 1]]])
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return color


def _random_spd(n, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=6 / n, random_state=seed + 1, format="csr")
    A = abs(A + A.T)
    A = A + sp.diags(np.asarray(A.sum(axis=1)).ravel() + rng.uniform(0.5, 1.5, n))
    return sp.csr_matrix(A)


def test_fsai_factor_of_a_block_diagonal_batch_is_per_block(driver, tmp_path):
    """A block-diagonal batch in the solver's batch order (colour-major, each
    colour cut into one chunk per system): the factor is block-diagonal and
    each block equals, bit for bit, the factor of that system alone in its
    own colour-major order (a row's lower-colour neighbours never leave its
    system and keep their relative order, so every local system is the
    same)."""
    systems = [_random_spd(n, 20 + k) for k, n in enumerate((150, 90, 220))]
    colours = [_greedy_colours(A) for A in systems]
    # each system alone, colour-major
    alone = []
    for A, col in zip(systems, colours):
        perm = np.argsort(col, kind="stable")
        Mp = sp.csr_matrix(A[perm][:, perm])
        Mp.sort_indices()
        rs, spl, cidx, val, dg = _split(Mp)
        g, gd = _run(driver, tmp_path, Mp.shape[0], rs, spl, cidx, val, dg)
        alone.append((perm, Mp, rs, spl, cidx, g, gd))
    # the batch: colour c of system 0, colour c of system 1, ..., then colour c + 1
    offsets = np.cumsum([0] + [A.shape[0] for A in systems])
    N = offsets[-1]
    where = {}  # (system, local colour-major row) -> batch row
    at = 0
    for c in range(max(col.max() for col in colours) + 1):
        for s, (perm, *_rest) in enumerate(alone):
            for li in np.flatnonzero(colours[s][perm] == c):
                where[(s, li)] = at
                at += 1
    assert at == N
    rows, cols, vals = [], [], []
    for s, (perm, Mp, *_rest) in enumerate(alone):
        C = Mp.tocoo()
        rows += [where[(s, i)] for i in C.row]
        cols += [where[(s, j)] for j in C.col]
        vals += list(C.data)
    B = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    B.sort_indices()
    rs, spl, cidx, val, dg = _split(B)
    g, gd = _run(driver, tmp_path, N, rs, spl, cidx, val, dg)
    for s, (perm, Mp, ars, aspl, acol, ag, agd) in enumerate(alone):
        back = {where[(s, i)]: i for i in range(Mp.shape[0])}
        for i in range(Mp.shape[0]):
            bi = where[(s, i)]
            assert gd[bi] == agd[i]
            # the batch row's entries, in order, map to the lone row's
            bcols = [back[c] for c in cidx[rs[bi]:rs[bi + 1]]]
            assert bcols == list(acol[ars[i]:ars[i + 1]])
            assert spl[bi] - rs[bi] == aspl[i] - ars[i]
            assert np.array_equal(g[rs[bi]:rs[bi + 1]], ag[ars[i]:ars[i + 1]])
