"""LAFEM file formats (fm_csr, fm_dv, fm_mtx, fm_exp) of the host layer.

Bar: files written are byte-identical to what the reference's own write_out
writes for the same arrays (oracle/_ref where it is built, else the fixtures under
tests/golden/io written by it); files read give the arrays the reference reads,
exactly for the binary images and the text files' 7 significant digits.
"""
import os

import numpy as np
import pytest

from oracle.pyoracle import Csr
from tests.golden.make_io_fixtures import fixture_matrix, fixture_symmetric, fixture_vector
from tests.helpers import random_csr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
FM = {"exp": 0, "dv": 1, "mtx": 2, "csr": 4, "binary": 13}


def read_bytes(p):
    with open(p, "rb") as f:
        return f.read()


def same_csr(a, b):
    assert (int(a["rows"]), int(a["cols"])) == (int(b.rows), int(b.cols))
    assert np.array_equal(a["rowptr"], np.asarray(b.rowptr, np.uint64))
    assert np.array_equal(a["col"], np.asarray(b.col, np.uint64))
    assert np.array_equal(a["val"].view(np.uint64), np.asarray(b.val, np.float64).view(np.uint64))


def as_text_values(v):
    """What a %.6e text file holds for v."""
    return np.array([float(f"{x:.6e}") for x in np.asarray(v, np.float64)])


# ---------------- committed fixtures written by the reference ----------------

@pytest.mark.parametrize("name,mode,fix,sym", [("matrix.csr", "csr", fixture_matrix, False),
                                               ("matrix_general.mtx", "mtx", fixture_matrix, False),
                                               ("symmetric.csr", "csr", fixture_symmetric, False),
                                               ("symmetric_lower.mtx", "mtx", fixture_symmetric, True)])
def test_matrix_fixtures(feat, tmp_path, name, mode, fix, sym):
    A = fix()
    got = feat.csr_read_file(mode, os.path.join(GOLD, name))
    if mode == "csr":
        same_csr(got, A)
    else:
        B = Csr(A.rows, A.cols, as_text_values(A.val), A.col, A.rowptr)
        same_csr(got, B)
    out = tmp_path / name
    feat.csr_write_file(mode, out, A.rows, A.cols, A.val, A.col, A.rowptr, symmetric=sym)
    assert read_bytes(out) == read_bytes(os.path.join(GOLD, name))


@pytest.mark.parametrize("name,mode", [("vector.dv", "dv"), ("vector.mtx", "mtx"), ("vector.exp", "exp")])
def test_vector_fixtures(feat, tmp_path, name, mode):
    v = fixture_vector()
    got = feat.dv_read_file(mode, os.path.join(GOLD, name))
    want = v if mode == "dv" else as_text_values(v)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    out = tmp_path / name
    feat.dv_write_file(mode, out, v)
    assert read_bytes(out) == read_bytes(os.path.join(GOLD, name))


def test_binary_mode_alias(feat, tmp_path):
    """fm_binary reads and writes the container's own image (fm_csr / fm_dv)."""
    A = fixture_matrix()
    feat.csr_write_file("binary", tmp_path / "a.bin", A.rows, A.cols, A.val, A.col, A.rowptr)
    assert read_bytes(tmp_path / "a.bin") == read_bytes(os.path.join(GOLD, "matrix.csr"))
    same_csr(feat.csr_read_file("binary", os.path.join(GOLD, "matrix.csr")), A)
    assert np.array_equal(feat.dv_read_file("binary", os.path.join(GOLD, "vector.dv")), fixture_vector())


# ---------------- against the reference build, random inputs ----------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_csr_files_match_reference(feat, reference, tmp_path, seed):
    rng = np.random.default_rng(seed)
    A = random_csr(rng, 40, 33, min_len=1, max_len=12, long_rows=[(5, 33)])
    for mode, sym in (("csr", False), ("mtx", False)):
        ours, theirs = tmp_path / f"o.{mode}", tmp_path / f"r.{mode}"
        feat.csr_write_file(mode, ours, A.rows, A.cols, A.val, A.col, A.rowptr, symmetric=sym)
        reference.csr_write(FM[mode], theirs, A, symmetric=sym)
        assert read_bytes(ours) == read_bytes(theirs), mode
        R = reference.csr_read(FM[mode], ours)
        got = feat.csr_read_file(mode, theirs)
        same_csr(got, R)
        if mode == "csr":
            same_csr(got, A)


def test_symmetric_mtx_matches_reference(feat, reference, tmp_path):
    rng = np.random.default_rng(11)
    n = 30
    M = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.2)
    M = M + M.T + np.diag(np.full(n, 5.0))
    rp, col, val = [0], [], []
    for r in range(n):
        nz = np.nonzero(M[r])[0]
        col.extend(nz.tolist())
        val.extend(M[r, nz].tolist())
        rp.append(len(col))
    A = Csr(n, n, np.array(val), np.array(col, np.uint64), np.array(rp, np.uint64))
    ours, theirs = tmp_path / "o.mtx", tmp_path / "r.mtx"
    feat.csr_write_file("mtx", ours, n, n, A.val, A.col, A.rowptr, symmetric=True)
    reference.csr_write(FM["mtx"], theirs, A, symmetric=True)
    assert read_bytes(ours) == read_bytes(theirs)
    # reading the lower triangle back gives the whole matrix, as the reference's reader does
    got = feat.csr_read_file("mtx", theirs)
    same_csr(got, reference.csr_read(FM["mtx"], theirs))
    same_csr(got, Csr(n, n, as_text_values(A.val), A.col, A.rowptr))


@pytest.mark.parametrize("mode", ["dv", "mtx", "exp"])
def test_vector_files_match_reference(feat, reference, tmp_path, mode):
    rng = np.random.default_rng(5)
    v = rng.standard_normal(101) * 10.0 ** rng.integers(-30, 30, 101)
    v[3], v[7] = 0.0, -0.0
    ours, theirs = tmp_path / "o", tmp_path / "r"
    feat.dv_write_file(mode, ours, v)
    reference.dv_write(FM[mode], theirs, v)
    assert read_bytes(ours) == read_bytes(theirs)
    got = feat.dv_read_file(mode, theirs)
    assert np.array_equal(got.view(np.uint64), reference.dv_read(FM[mode], ours).view(np.uint64))


# ---------------- semantics the reference leaves open, and error handling ----------------

def test_mtx_empty_rows_and_duplicates(feat, tmp_path):
    """Rows without entries keep correct row pointers; a repeated position keeps its
    first value (the reference's std::map insert) and is stored once."""
    p = tmp_path / "e.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n4 3 5\n"
                 "4 3 1.5\n1 2 -2.0\n4 1 3.0\n1 2 9.0\n4 3 7.0\n")
    got = feat.csr_read_file("mtx", p)
    assert got["rowptr"].tolist() == [0, 1, 1, 1, 3]
    assert got["col"].tolist() == [1, 0, 2]
    assert got["val"].tolist() == [-2.0, 3.0, 1.5]


def test_csr_roundtrip_exact_with_empty_rows(feat, tmp_path):
    rng = np.random.default_rng(3)
    A = random_csr(rng, 50, 20, min_len=0, max_len=6)
    feat.csr_write_file("csr", tmp_path / "a.csr", A.rows, A.cols, A.val, A.col, A.rowptr)
    same_csr(feat.csr_read_file("csr", tmp_path / "a.csr"), A)


@pytest.mark.parametrize("content,mode,kind", [
    ("%%MatrixMarket matrix array real general\n2 1\n1\n2\n", "mtx", "csr"),       # vector header
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "mtx", "csr"),   # row out of range
    ("%%MatrixMarket matrix array real general\n3 1\n1\n2\n", "mtx", "dv"),        # too few values
    ("%%MatrixMarket matrix array real general\n3 2\n1\n2\n", "mtx", "dv"),        # not n x 1
])
def test_malformed_text_files_raise(feat, tmp_path, content, mode, kind):
    p = tmp_path / "bad"
    p.write_text(content)
    with pytest.raises(feat.FeatError):
        (feat.csr_read_file if kind == "csr" else feat.dv_read_file)(mode, p)


def test_malformed_binary_files_raise(feat, tmp_path):
    good = read_bytes(os.path.join(GOLD, "matrix.csr"))
    cases = {
        "truncated": good[:200],
        "wrong_mode": good[:8] + (1).to_bytes(8, "little") + good[16:],
        "compressed": good[:80] + (0x44).to_bytes(8, "little") + good[88:],
        "float_values": good[:16] + (0x0000000600000004).to_bytes(8, "little") + good[24:],
    }
    for name, data in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        with pytest.raises(feat.FeatError):
            feat.csr_read_file("csr", p)
    with pytest.raises(feat.FeatError):          # a vector image is not a matrix image
        feat.csr_read_file("csr", os.path.join(GOLD, "vector.dv"))
    with pytest.raises(feat.FeatError):
        feat.csr_read_file("exp", os.path.join(GOLD, "matrix.csr"))
    with pytest.raises(feat.FeatError):
        feat.dv_read_file("dv", tmp_path / "missing.dv")
