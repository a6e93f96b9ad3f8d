"""CPU restatement of the previous-solutions projection (ofeli_amd/csrc/ofh_proj.hip).

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels; only tests/ use it.
No reference counterpart (the reference starts from zero or from the previous
solution, TimeStepping.cpp:207).  Definition (Fischer 1998): a basis V (at
most M directions) of the span of recent solutions, A-orthonormal (guess
x0 = V V^T b, the A-norm-closest vector of the span) or, minres=True, with AV
orthonormal (x0 = V (AV)^T b, the least-residual vector).  Update with a
solution x: two classical Gram-Schmidt passes in the mode's inner product; a
direction with squared norm <= 1e-12 that of x is dropped (x is already in
the span up to residual noise); otherwise it is normalised and appended and
x joins the ring of the last M accepted solutions.  When V is full, an
accepted x rebuilds it from the last M/2 - 1 ring solutions and x.
"""
import numpy as np


class Projection:
    def __init__(self, A, M, minres=False):
        self.A = A
        self.M = M
        self.minres = minres
        n = A.shape[0]
        self.V = np.zeros((M, n))
        self.AV = np.zeros((M, n))
        self.head = 0
        self.count = 0

    def guess(self, b):
        return ((self.AV if self.minres else self.V) @ b) @ self.V

    def update(self, x):
        x = np.array(x, dtype=np.float64)
        if not hasattr(self, "X"):
            self.X = []
        if self.count < self.M:
            acc = self._insert(x)
        else:
            acc = self._insert(x, dry=True)
            if acc:
                self.V[:] = 0.0
                self.AV[:] = 0.0
                self.head = self.count = 0
                for y in self.X[-(self.M // 2 - 1):]:
                    self._insert(y)
                self._insert(x)
        if acc:
            self.X = (self.X + [x])[-self.M:]
        return acc

    def _insert(self, x, dry=False):
        Ax = self.A @ x
        mr = self.minres
        c = self.AV @ (Ax if mr else x)
        d = x - c @ self.V
        Ad = Ax - c @ self.AV
        xax = float(Ax @ Ax) if mr else float(x @ Ax)
        c2 = self.AV @ (Ad if mr else d)
        d = d - c2 @ self.V
        Ad = Ad - c2 @ self.AV
        dad = float(Ad @ Ad) if mr else float(d @ Ad)
        ok = dad > 1e-12 * xax and dad > 0
        if ok and not dry:
            s = 1.0 / np.sqrt(dad)
            self.V[self.head] = d * s
            self.AV[self.head] = Ad * s
            self.head += 1
            self.count += 1
        return ok
