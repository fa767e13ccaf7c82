"""Restatement of the collision-matrix ingest (Fill_M_Bin!, GainMatrix_to_M_Bin!,
LossMatrix_to_M_Bin!; src/Collisions/BuildBinaryMatrices.jl:81-284) as plain Python loops that
write the monolithic M_Bin (N^2 x N, row (b-1) N + a, column c) exactly as the reference does:
dense (`M[r, c] += v`) or sparse (the pushed (I, J, V) triples, summed per entry in push order as
sparse(I, J, V, m, n) does, BuildBinaryMatrices.jl:48).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker for dtx_add_binary_interaction.  The
loops run over every array entry, so they are for tiny grids (N of a few tens).
"""
from __future__ import annotations

import numpy as np

ANI, AXI, ISO = 0, 1, 2


def _vsum(v):
    """sum(v) of a short vector, left to right."""
    s = 0.0
    for x in v:
        s = s + float(x)
    return s


class _Writer:
    """M_Bin as the reference fills it: dense accumulation or pushed (row, col, value)."""

    def __init__(self, N, sparse):
        self.N, self.sparse = N, sparse
        self.M = np.zeros((N * N, N), order="F")
        self.I, self.J, self.V = [], [], []

    def add(self, row, col, v):  # 1-based row / col, as the reference's indices
        if self.sparse:
            self.I.append(row)
            self.J.append(col)
            self.V.append(v)
        else:
            self.M[row - 1, col - 1] = self.M[row - 1, col - 1] + v

    def result(self):
        """The dense N^2 x N array (the sparse triples summed per entry in push order)."""
        if not self.sparse:
            return self.M
        M = np.zeros((self.N * self.N, self.N), order="F")
        seen = np.zeros(M.shape, dtype=bool)
        for r, c, v in zip(self.I, self.J, self.V):
            if seen[r - 1, c - 1]:
                M[r - 1, c - 1] = M[r - 1, c - 1] + v
            else:
                M[r - 1, c - 1] = v
                seen[r - 1, c - 1] = True
        return M


def gain_to_m_bin(W, G, off3, off1, off2, mode, dpy1, dpz1, dpy2, dpz2, dpy3, dpz3):
    """GainMatrix_to_M_Bin! (BuildBinaryMatrices.jl:114-204); offsets 0-based, indices below
    1-based as written there."""
    N = W.N
    px3_num = G.shape[0] - 2
    py3_num, pz3_num, px1_num, py1_num, pz1_num, px2_num, py2_num, pz2_num = G.shape[1:]
    is_sparse = W.sparse
    for px2 in range(1, px2_num + 1):
        for px1 in range(1, px1_num + 1):
            for px3 in range(1, px3_num + 1):
                if mode == ISO:
                    val = 0.0
                    w = 1.0 / (_vsum(dpz1) * _vsum(dpz2) * _vsum(dpz3) * _vsum(dpy1) *
                               _vsum(dpy2) * _vsum(dpy3))
                    for py1 in range(1, py1_num + 1):
                        for pz1 in range(1, pz1_num + 1):
                            for py2 in range(1, py2_num + 1):
                                for pz2 in range(1, pz2_num + 1):
                                    for py3 in range(1, py3_num + 1):
                                        for pz3 in range(1, pz3_num + 1):
                                            val += (G[px3, py3 - 1, pz3 - 1, px1 - 1, py1 - 1,
                                                      pz1 - 1, px2 - 1, py2 - 1, pz2 - 1] *
                                                    dpz1[pz1 - 1] * dpz2[pz2 - 1] *
                                                    dpz3[pz3 - 1] * dpy1[py1 - 1] *
                                                    dpy2[py2 - 1] * dpy3[py3 - 1])
                for py1 in range(1, py1_num + 1):
                    for py2 in range(1, py2_num + 1):
                        for py3 in range(1, py3_num + 1):
                            if mode == AXI:
                                val = 0.0
                                w = 1.0 / (_vsum(dpz1) * _vsum(dpz2) * _vsum(dpz3))
                                for pz1 in range(1, pz1_num + 1):
                                    for pz2 in range(1, pz2_num + 1):
                                        for pz3 in range(1, pz3_num + 1):
                                            val += (G[px3, py3 - 1, pz3 - 1, px1 - 1, py1 - 1,
                                                      pz1 - 1, px2 - 1, py2 - 1, pz2 - 1] *
                                                    dpz1[pz1 - 1] * dpz2[pz2 - 1] *
                                                    dpz3[pz3 - 1])
                            for pz1 in range(1, pz1_num + 1):
                                for pz2 in range(1, pz2_num + 1):
                                    for pz3 in range(1, pz3_num + 1):
                                        if mode == ANI:
                                            col = G[:, py3 - 1, pz3 - 1, px1 - 1, py1 - 1,
                                                    pz1 - 1, px2 - 1, py2 - 1, pz2 - 1]
                                            gmax = float(np.max(col)) if is_sparse else None
                                            val = float(col[px3])
                                            w = 1.0
                                            if is_sparse and val < gmax * 1e-16:
                                                continue
                                        if val == 0.0:
                                            continue
                                        a = ((pz3 - 1) * px3_num * py3_num + (py3 - 1) * px3_num
                                             + px3 + off3)
                                        b = ((pz1 - 1) * px1_num * py1_num + (py1 - 1) * px1_num
                                             + px1 + off1)
                                        c = ((pz2 - 1) * px2_num * py2_num + (py2 - 1) * px2_num
                                             + px2 + off2)
                                        v = val * w / 2
                                        W.add((b - 1) * N + (a - 1) + 1, c, v)
                                        W.add((c - 1) * N + (a - 1) + 1, b, v)


def loss_to_m_bin(W, L, off1, off2, mode, dpy1, dpz1, dpy2, dpz2):
    """LossMatrix_to_M_Bin! (BuildBinaryMatrices.jl:206-284)."""
    N = W.N
    px1_num, py1_num, pz1_num, px2_num, py2_num, pz2_num = L.shape
    for px2 in range(1, px2_num + 1):
        for px1 in range(1, px1_num + 1):
            if mode == ISO:
                val = 0.0
                w = 1.0 / (_vsum(dpz1) * _vsum(dpz2) * _vsum(dpy1) * _vsum(dpy2))
                for py2 in range(1, py2_num + 1):
                    for py1 in range(1, py1_num + 1):
                        for pz2 in range(1, pz2_num + 1):
                            for pz1 in range(1, pz1_num + 1):
                                val += (L[px1 - 1, py1 - 1, pz1 - 1, px2 - 1, py2 - 1, pz2 - 1] *
                                        dpz1[pz1 - 1] * dpz2[pz2 - 1] * dpy1[py1 - 1] *
                                        dpy2[py2 - 1])
            for py2 in range(1, py2_num + 1):
                for py1 in range(1, py1_num + 1):
                    if mode == AXI:
                        val = 0.0
                        w = 1.0 / (_vsum(dpz1) * _vsum(dpz2))
                        for pz2 in range(1, pz2_num + 1):
                            for pz1 in range(1, pz1_num + 1):
                                val += (L[px1 - 1, py1 - 1, pz1 - 1, px2 - 1, py2 - 1, pz2 - 1] *
                                        dpz1[pz1 - 1] * dpz2[pz2 - 1])
                    for pz2 in range(1, pz2_num + 1):
                        for pz1 in range(1, pz1_num + 1):
                            if mode == ANI:
                                val = float(L[px1 - 1, py1 - 1, pz1 - 1, px2 - 1, py2 - 1,
                                              pz2 - 1])
                                w = 1.0
                            if val == 0.0:
                                continue
                            a = (pz1 - 1) * px1_num * py1_num + (py1 - 1) * px1_num + px1 + off1
                            b = a
                            c = (pz2 - 1) * px2_num * py2_num + (py2 - 1) * px2_num + px2 + off2
                            v = -(val * w / 2)
                            W.add((b - 1) * N + (a - 1) + 1, c, v)
                            W.add((c - 1) * N + (a - 1) + 1, b, v)


def fill_m_bin(N, offsets, dpy, dpz, interactions, mode=ANI, sparse=False):
    """M_Bin (N^2 x N, Fortran) after Fill_M_Bin! (BuildBinaryMatrices.jl:81-112) of each
    interaction in order.  offsets / dpy / dpz: per species; interactions: tuples
    (name_locs 0-based, GainMatrix3, GainMatrix4, LossMatrix1, LossMatrix2, gain_scale,
    loss_scale), the arrays unscaled (LoadMatrices.jl:129-133 multiplies them first)."""
    W = _Writer(N, sparse)
    for locs, G3, G4, L1, L2, gs, ls in interactions:
        n1, n2, n3, n4 = locs
        G3, G4 = np.asarray(G3) * gs, np.asarray(G4) * gs
        L1, L2 = np.asarray(L1) * ls, np.asarray(L2) * ls
        gain_to_m_bin(W, G3, offsets[n3], offsets[n1], offsets[n2], mode, dpy[n1], dpz[n1],
                      dpy[n2], dpz[n2], dpy[n3], dpz[n3])
        gain_to_m_bin(W, G4, offsets[n4], offsets[n1], offsets[n2], mode, dpy[n1], dpz[n1],
                      dpy[n2], dpz[n2], dpy[n4], dpz[n4])
        loss_to_m_bin(W, L1, offsets[n1], offsets[n2], mode, dpy[n1], dpz[n1], dpy[n2], dpz[n2])
        loss_to_m_bin(W, L2, offsets[n2], offsets[n1], mode, dpy[n2], dpz[n2], dpy[n1], dpz[n1])
    return W.result()
