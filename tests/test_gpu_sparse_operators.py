"""Sparse operators at the boundary (Bin_sparse / Emi_sparse): M_Bin (N^2 x N) and M_Emi (n x n)
handed over as SparseMatrixCSC{Float64,Int64} through dtx_set_binary_csc / dtx_set_emission_csc
give the dense path's states bit for bit, and the oracle's (with its sparse M_Emi) at 1e-12;
stored zeros are harmless; an emission entry outside the per-cell blocks is refused."""
import numpy as np
import pytest

import oracle as O
from dtx import _lib, workloads
from dtx.api import Solve

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def rel_linf(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("method", ["euler", "symplectic"])
@pytest.mark.parametrize("kw", [dict(emission=True, injection=True),
                                dict(num_species=3, x_num=2, emission=True),
                                dict(domain=[0, 2], emission=True)],
                         ids=["2species", "3species", "domain"])
def test_sparse_equals_dense_and_oracle(kw, method):
    w = workloads.tiny(steps=20, **kw)
    ts = w.t_save(3)
    dense = workloads.make_struct(w, method=method)
    sparse = workloads.make_struct(w, method=method, sparse=True)
    od = Solve(dense, w.dt, ts, Verbose=0)
    osp = Solve(sparse, w.dt, ts, Verbose=0)
    for a, b in zip(osp.f, od.f):
        np.testing.assert_array_equal(a, b)
    fo, *_ = O.problem_from_workload(w, sparse_emission=True).solve(
        w.f0, w.dt, ts, verbose=0, method=0 if method == "euler" else 1)
    for i in range(len(ts)):
        assert rel_linf(osp.f[i], fo[i]) < TOL64, i
    dense.close()
    sparse.close()


def test_stored_zeros_in_sparse_binary():
    """Entries stored with value 0 (which Julia's sparse() keeps when they are pushed) change
    nothing: the same as the dense tensor with those entries zero."""
    w = workloads.tiny(steps=10)
    colptr, rowval, nzval = workloads.binary_csc(w)
    nz = nzval.copy()
    nz[::7] = 0.0
    N = w.ps.n_momentum
    M = np.asarray(w.M_Bin).reshape(N * N, N, order="F").copy()
    for j in range(N):
        for q in range(colptr[j] - 1, colptr[j + 1] - 1):
            M[rowval[q] - 1, j] = nz[q]
    w_dense = workloads.tiny(steps=10)
    w_dense.M_Bin = M.reshape((N, N, N), order="F")
    w.M_Bin = None
    from dtx.api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                         ForwardEulerStruct)
    sparse = ForwardEulerStruct(w.ps, w.f0, w.inj,
                                BinaryMatricesStruct(M_Bin=(colptr, rowval, nz),
                                                     Binary_list=["sparse"]),
                                EmissionMatricesStruct(), FluxMatricesStruct(w.Ap, w.X, w.P, w.Vol))
    dense = workloads.make_struct(w_dense)
    ts = w_dense.t_save(2)
    a = Solve(sparse, w_dense.dt, ts, Verbose=0).f[-1]
    b = Solve(dense, w_dense.dt, ts, Verbose=0).f[-1]
    np.testing.assert_array_equal(a, b)
    sparse.close()
    dense.close()


def test_emission_outside_blocks_refused():
    w = workloads.tiny(steps=2, emission=True)
    colptr, rowval, nzval = workloads.emission_csc(w)
    N = w.ps.n_momentum
    # move the first stored entry of column 0 into the second cell's rows
    rowval = rowval.copy()
    rowval[0] = N + 1
    rowval[:N] = np.sort(rowval[:N])
    with pytest.raises(_lib.DtxError, match="outside the per-cell diagonal blocks"):
        from dtx.api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                             ForwardEulerStruct)
        ForwardEulerStruct(w.ps, w.f0, w.inj, BinaryMatricesStruct(),
                           EmissionMatricesStruct(M_Emi=(colptr, rowval, nzval),
                                                  Emission_list=["sparse"]),
                           FluxMatricesStruct(w.Ap, w.X, w.P, w.Vol))
