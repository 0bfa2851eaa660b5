"""Matrix Market I/O of the C++ API (include/lal_linalg.hpp, lal_linalg::mtx)
against LinAlgLib's own csr_matrix::read_mtx / write_mtx (compiled from its
sources into oracle/_ref): the same CSR arrays bit for bit from the same
files, the same refusals, and byte-identical files written from the same
arrays.  Host only; runs without a GPU."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import CSR, Reference

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "tests", "cpp", "mtx_tool")

pytestmark = pytest.mark.skipif(not os.path.exists(TOOL),
                                reason="tests/cpp/mtx_tool not built (python -m linalglib_amd.build)")


def ours_read(path, tmp_path):
    out = tmp_path / "parsed.bin"
    r = subprocess.run([TOOL, "read", str(path), str(out)], capture_output=True, text=True)
    if r.returncode == 3:
        return None
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    m, n, nnz = np.frombuffer(raw[:12], dtype=np.int32)
    o = 12
    rp = np.frombuffer(raw[o: o + 4 * (m + 1)], dtype=np.int32)
    o += 4 * (m + 1)
    ci = np.frombuffer(raw[o: o + 4 * nnz], dtype=np.int32)
    o += 4 * nnz
    v = np.frombuffer(raw[o: o + 8 * nnz], dtype=np.float64)
    return CSR(rp, ci, v, int(m), int(n))


def ours_write(A, path, tmp_path):
    blob = tmp_path / "arrays.bin"
    blob.write_bytes(np.array([A.m, A.n, A.nnz], np.int32).tobytes() + A.row_ptr.tobytes()
                     + A.col_ind.tobytes() + A.val.tobytes())
    r = subprocess.run([TOOL, "write", str(blob), str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def same(a, b):
    return (a.m == b.m and a.n == b.n and np.array_equal(a.row_ptr, b.row_ptr)
            and np.array_equal(a.col_ind, b.col_ind)
            and np.array_equal(a.val.view(np.int64), b.val.view(np.int64)))


def write_text(path, header, m, n, entries, comments=("% a comment",)):
    lines = [header] + list(comments) + [f"{m} {n} {len(entries)}"]
    lines += [f"{r} {c} {v}" for r, c, v in entries]
    path.write_text("\n".join(lines) + "\n")


def cases(rng):
    # general real, unsorted, with a duplicate pair
    ent = [(3, 1, 0.5), (1, 2, -1.25), (1, 1, 4.0), (2, 3, 1e-7), (3, 1, 0.25), (2, 2, 3.0)]
    yield "general", "%%MatrixMarket matrix coordinate real general", 3, 4, ent
    # symmetric: lower triangle given, diagonal once
    m = 30
    ent = []
    for i in range(1, m + 1):
        ent.append((i, i, float(rng.uniform(3, 5))))
        for j in rng.choice(np.arange(1, i), size=min(i - 1, 3), replace=False) if i > 1 else []:
            ent.append((i, int(j), float(rng.uniform(-1, 0))))
    rng.shuffle(ent)
    yield "symmetric", "%%MatrixMarket matrix coordinate real symmetric", m, m, ent
    # integer field
    yield ("integer", "%%MatrixMarket matrix coordinate integer general", 4, 4,
           [(1, 1, 7), (4, 4, -3), (2, 3, 12), (3, 2, 5)])
    # random rectangular with empty rows and repeated values printed in scientific form
    ent = [(int(r), int(c), float(v)) for r, c, v in
           zip(rng.integers(1, 41, 120), rng.integers(1, 26, 120), rng.standard_normal(120))]
    yield "random", "%%MatrixMarket matrix coordinate real general", 50, 25, ent
    # empty matrix
    yield "empty", "%%MatrixMarket matrix coordinate real general", 5, 5, []


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")
def test_read_matches_reference(tmp_path):
    ref = Reference()
    rng = np.random.default_rng(3)
    for name, header, m, n, ent in cases(rng):
        p = tmp_path / f"{name}.mtx"
        write_text(p, header, m, n, ent)
        a = ours_read(p, tmp_path)
        b = ref.read_mtx(p)
        assert a is not None and b is not None, name
        assert same(a, b), name
        # columns sorted within rows, duplicates summed
        for i in range(a.m):
            assert np.all(np.diff(a.col_ind[a.row_ptr[i]: a.row_ptr[i + 1]]) > 0)


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")
def test_refusals_match_reference(tmp_path):
    ref = Reference()
    for header in ("%%MatrixMarket matrix array real general",
                   "%%MatrixMarket matrix coordinate complex general",
                   "%%MatrixMarket matrix coordinate pattern general",
                   "%%MatrixMarket matrix coordinate real hermitian",
                   "%%MatrixMarket matrix coordinate real skew-symmetric",
                   "%%MatrixMarket vector coordinate real general"):
        p = tmp_path / "x.mtx"
        write_text(p, header, 2, 2, [(1, 1, 1.0), (2, 2, 2.0)])
        assert ours_read(p, tmp_path) is None, header
        assert ref.read_mtx(p) is None, header


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")
def test_write_matches_reference_bytes(tmp_path):
    ref = Reference()
    rng = np.random.default_rng(5)
    from linalglib_amd import problems as pb

    rp, ci, v, m = pb.poisson2d(6)
    mats = [CSR(rp, ci, v, m, m)]
    rp = np.array([0, 2, 2, 5], np.int32)
    mats.append(CSR(rp, np.array([0, 3, 1, 2, 4], np.int32),
                    np.array([1.0 / 3, -2.5e-11, 12345.678901234, -0.5, 7.0]), 3, 5))
    mats.append(CSR(np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0), 3, 3))
    for k, A in enumerate(mats):
        p1, p2 = tmp_path / f"ours{k}.mtx", tmp_path / f"ref{k}.mtx"
        ours_write(A, p1, tmp_path)
        assert ref.write_mtx(p2, A)
        assert p1.read_bytes() == p2.read_bytes(), k


def test_round_trip_closed_form(tmp_path):
    """write then read returns the matrix (values with 10 decimals: exact for
    the Poisson stencil), without the reference."""
    from linalglib_amd import problems as pb

    rp, ci, v, m = pb.poisson3d(5)
    A = CSR(rp, ci, v, m, m)
    p = tmp_path / "p.mtx"
    ours_write(A, p, tmp_path)
    B = ours_read(p, tmp_path)
    assert same(A, B)
