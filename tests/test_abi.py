"""CPU checks of the C ABI and the host-side logic: the library loads, exports
every entry point include/lal.h declares, and the host-only helpers behave.
No test here calls into the GPU."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lal.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s+(lal_\w+)\s*\(", text,
                                 flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(lal.lib_path()):
        from linalglib_amd import build

        build.build()
    return lal.load()


def test_every_declared_symbol_is_exported(lib):
    names = declared_symbols()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_defaults(lib):
    assert lib.lal_status_string(0) == b"ok"
    assert lib.lal_status_string(1) == b"invalid argument"
    assert lib.lal_status_string(99) == b"unknown status"
    o = lal.saamg_options()
    assert (o.max_levels, o.eps, o.relax) == (20, 0.08, 2.0 / 3.0)
    s = lal.solve_options()
    assert (s.rel_tol, s.abs_tol, s.max_iter) == (1e-6, 1e-6, 1000)  # iter_control defaults


def test_null_arguments_are_refused(lib):
    # argument checks run before any HIP call
    assert lib.lal_csrmv_analysis(None, 1, 1, 1, None, None, 0, None) == 1
    assert lib.lal_axpy_f64(None, 4, 1.0, None, None) == 1
    assert lib.lal_saamg_setup(None, 4, 4, 4, None, None, None, None, None) == 1


def test_handle_refuses_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(lal.LalError):
        lal.Handle()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_byte_formulas():
    sys.path.insert(0, ROOT)
    import bench

    m, nnz = 2097152, 14581760
    assert bench.spmv_bytes(m, m, nnz, 8, 8) == 216_924_164  # 8(nnz+m+n)+4(m+1+nnz)
    assert bench.spmv_bytes(m, m, nnz, 4, 4) == 141_819_908
    # 16-bit columns: 2 bytes per nonzero less
    assert bench.spmv_bytes(m, m, nnz, 8, 8, cbytes=2) == 216_924_164 - 2 * nnz
    # one-level "hierarchy": only the PCG terms remain
    assert bench.solve_bytes([(10, 30, 0, 4, 4, 4)], 1, 8) == \
        bench.spmv_bytes(10, 10, 30, 8, 8) + 14 * 8 * 10
    # two levels (A0 10 x 10 with 30 nonzeros, P 10 x 4 with 12): the Krylov
    # SpMV, residual and post-sweep read A0's columns, the restriction R's,
    # the prolongation P's; each column width counts once per nonzero read
    lv = [(10, 30, 12, 2, 2, 4), (4, 10, 0, 4, 4, 4)]
    wide = [(10, 30, 12, 4, 4, 4), (4, 10, 0, 4, 4, 4)]
    assert bench.solve_bytes(wide, 1, 8) - bench.solve_bytes(lv, 1, 8) == \
        2 * 30 * 3 + 2 * 12  # SpMV, residual, post-sweep with A0; prolongation with P
    # one reliable update: an FP64 residual (SpMV bytes plus b read and r
    # written), the FP32 correction read with x read and written, y cleared,
    # the norm of r and its scaled FP32 copy
    upd = bench.reliable_update_bytes(m, nnz)
    assert upd == bench.spmv_bytes(m, m, nnz, 8, 8) + 8 * m + (4 + 16) * m + 4 * m + 8 * m + 12 * m
    assert bench.reliable_update_bytes(m, nnz, 2) == upd - 2 * nnz


def test_poisson_generators_symmetric_and_sorted():
    for rp, ci, v, m in (pb.poisson2d(7, 5), pb.poisson3d(4, 5, 6)):
        A = CSR(rp, ci, v, m, m).to_scipy()
        assert (A != A.T).nnz == 0
        for i in range(m):
            assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)


def test_diffusion_generator_row_sums():
    # interior rows sum to 0 (conservative fluxes); Dirichlet faces add their
    # conductance to the diagonal only
    k = np.exp(np.random.default_rng(0).standard_normal(5 ** 3))
    rp, ci, v, m = pb.diffusion3d(5, cell_kappa=k)
    A = CSR(rp, ci, v, m, m).to_scipy()
    assert abs(A - A.T).max() == 0
    rs = np.asarray(A.sum(axis=1)).ravel().reshape(5, 5, 5)
    assert np.allclose(rs[1:-1, 1:-1, 1:-1], 0, atol=1e-12)
    assert np.all(rs > -1e-12)
    assert np.all(np.linalg.eigvalsh(A.toarray()) > 0)
