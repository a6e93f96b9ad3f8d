"""CPU restatement of the previous-solutions projection (ofeli_amd/csrc/ofh_proj.hip).

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels; only tests/ use it.
No reference counterpart (the reference starts from zero or from the previous
solution, TimeStepping.cpp:207).  Definition (Fischer 1998): an A-orthonormal
basis V (at most M directions) of the span of recent solutions; guess
x0 = V V^T b; update with a solution x: the solution is kept in a ring of the
last M solutions; when V is full it is rebuilt from the last M/2 solutions
(newest included), otherwise x is added: two classical Gram-Schmidt passes in
the A-inner product, A-normalisation; a direction with |d|_A^2 <= 1e-24
x^T A x is stored as zero.  The span always contains the newest solution, so
the guess is never worse than the warm start in the A-norm.
"""
import numpy as np


class Projection:
    def __init__(self, A, M):
        self.A = A
        self.M = M
        n = A.shape[0]
        self.V = np.zeros((M, n))
        self.AV = np.zeros((M, n))
        self.head = 0
        self.count = 0

    def guess(self, b):
        return (self.V @ b) @ self.V

    def update(self, x):
        if not hasattr(self, "X"):
            self.X = []
        self.X = (self.X + [np.array(x, dtype=np.float64)])[-self.M:]
        if self.count == self.M:
            self.V[:] = 0.0
            self.AV[:] = 0.0
            self.head = self.count = 0
            for y in self.X[-(self.M // 2):]:
                self._insert(y)
        else:
            self._insert(x)

    def _insert(self, x):
        Ax = self.A @ x
        c = self.AV @ x
        d = x - c @ self.V
        Ad = Ax - c @ self.AV
        xax = float(x @ Ax)
        c2 = self.AV @ d
        d = d - c2 @ self.V
        Ad = Ad - c2 @ self.AV
        dad = float(d @ Ad)
        s = 1.0 / np.sqrt(dad) if (dad > 1e-24 * xax and dad > 0) else 0.0
        self.V[self.head] = d * s
        self.AV[self.head] = Ad * s
        self.head = (self.head + 1) % self.M
        self.count = min(self.count + 1, self.M)
