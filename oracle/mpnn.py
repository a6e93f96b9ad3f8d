"""CPU forward/backward of the learned initial guess (message-passing network).

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels in
ofeli_amd/csrc/ofh_mpnn.hip.  Only tests/ (and bench.py's CPU leg) use it.

The network has no counterpart in the reference (which always starts a solve
from zero, Vect_impl.h:547, or from the previous step, TimeStepping.cpp:207);
it predicts the start vector x0 of the next solve A x = b in a sequence of
solves, from the last solution x_n and the one before, x_p.  Everything here
is numpy float64; with bf16=True the operands of the six dense products are
rounded to bfloat16 (round to nearest even) exactly where the GPU feeds its
matrix cores, so the GPU result is reproduced up to FP32 accumulation order.

Definition (n rows, A symmetric CSR with 0-based columns):
  x_b = x_n + ext (x_n - x_p)      (ext = 0: warm base, 1: linear extrapolation)
  r  = b - A x_b,  z = D^-1 r,  s = rms(z) (1 if 0),  z2 = D^-1 (r - A z)
  F  = [z/s, z2/s, (x_n - x_p)/s, ln a_ii - mean(ln a_ii), 0, 0, 0, 0]
  H  = relu(F W_enc + b_enc)
  per layer:  Z = H W_ab                    (64 -> 128)
              M_i = sum_{j in N(i), j != i} relu(Z_i[:64] + Z_j[64:] + e_ij w_e + b_m)
              U = relu([H, M] W_u + b_u)    (128 -> 64);  H = H + U
  d = s (H w_dec + b_dec)          (the learned search direction)
  x0 = x_b + alpha d,  alpha = d^T r / d^T A d  (0 if d^T A d <= 0)
with e_ij = (a_ij + a_ji) / (2 sqrt(a_ii a_jj)).  alpha is the exact line search in the
A-norm, so x0 is never worse than x_b.  Training loss (x* = the solver's
solution, e_b = x* - x_b, g = A e_b, E0 = e_b^T g):
  L = 1 - (d^T g)^2 / (d^T A d E0),
the A-norm error left after the line search relative to the base's (one
minus the squared A-cosine between d and e_b); it is invariant to the scale
of d.  dL/dd = -2 (d^T g) / (d^T A d E0) (g - (d^T g / d^T A d) A d).

Minimum-residual objective (obj = 1), aligned with CG's stop test on ||r||:
  x0 = x_b + alpha d,  alpha = (A d)^T r / (A d)^T A d    (0 if A d = 0)
  L = 1 - c^2 / (q R),  c = (A d)^T r, q = (A d)^T A d, R = r^T r,
the squared residual norm left after the line search relative to the base's;
it needs no reference solution.  With u = A d,
  dL/du = -2c/(qR) r + 2c^2/(q^2 R) u,   dL/dd = A dL/du   (A symmetric).
"""
import numpy as np
import scipy.sparse as sp

HID = 64
FIN = 8


def layer_size(h=HID):
    return h * 2 * h + h + h + 2 * h * h + h


def n_params(layers, h=HID):
    return FIN * h + h + layers * layer_size(h) + h + 1


def offsets(layers, h=HID):
    """Parameter offsets in the flat FP32 vector (same order as the GPU)."""
    o = {"enc_W": 0, "enc_b": FIN * h}
    p = FIN * h + h
    for l in range(layers):
        o["Wab%d" % l] = p
        o["we%d" % l] = p + 2 * h * h
        o["bm%d" % l] = p + 2 * h * h + h
        o["Wu%d" % l] = p + 2 * h * h + 2 * h
        o["bu%d" % l] = p + 4 * h * h + 2 * h
        p += layer_size(h)
    o["dec_w"] = p
    o["dec_b"] = p + h
    return o


def unpack(theta, layers, h=HID):
    o = offsets(layers, h)
    t = np.asarray(theta, np.float64)
    P = {"enc_W": t[o["enc_W"]:o["enc_W"] + FIN * h].reshape(FIN, h),
         "enc_b": t[o["enc_b"]:o["enc_b"] + h],
         "dec_w": t[o["dec_w"]:o["dec_w"] + h], "dec_b": t[o["dec_b"]]}
    for l in range(layers):
        P["Wab%d" % l] = t[o["Wab%d" % l]:o["Wab%d" % l] + 2 * h * h].reshape(h, 2 * h)
        P["we%d" % l] = t[o["we%d" % l]:o["we%d" % l] + h]
        P["bm%d" % l] = t[o["bm%d" % l]:o["bm%d" % l] + h]
        P["Wu%d" % l] = t[o["Wu%d" % l]:o["Wu%d" % l] + 2 * h * h].reshape(2 * h, h)
        P["bu%d" % l] = t[o["bu%d" % l]:o["bu%d" % l] + h]
    return P


def pack(G, layers, h=HID):
    o = offsets(layers, h)
    t = np.zeros(n_params(layers, h))
    for k, v in G.items():
        v = np.atleast_1d(np.asarray(v, np.float64)).ravel()
        t[o[k]:o[k] + v.size] = v
    return t


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """The matrix graph: CSR (0-based), diagonal, edge weights e_ij."""

    def __init__(self, row_ptr, col, val):
        self.n = len(row_ptr) - 1
        self.rp = np.asarray(row_ptr, np.int64)
        self.col = np.asarray(col, np.int64)
        self.val = np.asarray(val, np.float64)
        rows = np.repeat(np.arange(self.n), np.diff(self.rp))
        self.A = sp.csr_matrix((self.val, self.col, self.rp), shape=(self.n, self.n))
        diag = self.A.diagonal()
        self.dinv = 1.0 / diag
        off = self.col != rows
        self.ei, self.ej = rows[off], self.col[off]
        # e_ij of the symmetric part, in FP32 as the GPU stores it
        vt = self.A.T.tocsr()
        vt.sort_indices()
        valt = vt.data  # a_ji in the (sorted, symmetric) pattern of A
        self.e = ((self.val[off] + valt[off]) / (2.0 * np.sqrt(diag[self.ei] * diag[self.ej]))
                  ).astype(np.float32).astype(np.float64)
        ld = np.log(diag)
        self.ld = (ld - ld.sum() / self.n).astype(np.float32).astype(np.float64)
        self.S = sp.csr_matrix((np.ones(len(self.ei)), (self.ei, np.arange(len(self.ei)))),
                               shape=(self.n, len(self.ei)))
        self.St = sp.csr_matrix((np.ones(len(self.ej)), (self.ej, np.arange(len(self.ej)))),
                                shape=(self.n, len(self.ej)))


def base(x_n, x_p, ext=0):
    return x_n + float(ext) * (x_n - x_p)


def features(g, x_n, x_p, b, ext=0):
    r = b - g.A @ base(x_n, x_p, ext)
    z = g.dinv * r
    s = float(np.sqrt(np.dot(z, z) / g.n))
    if not s > 0:
        s = 1.0
    z2 = g.dinv * (r - g.A @ z)
    F = np.zeros((g.n, FIN))
    F[:, 0] = z / s
    F[:, 1] = z2 / s
    F[:, 2] = (x_n - x_p) / s
    F[:, 3] = g.ld
    return F.astype(np.float32).astype(np.float64), s


def forward(g, theta, layers, x_n, x_p, b, bf=False, ext=0, obj=0):
    """Returns x0 and the cache of the backward pass."""
    P = unpack(theta, layers)
    rnd = bf16 if bf else (lambda a: a)
    F, s = features(g, x_n, x_p, b, ext)
    H = np.maximum(F @ P["enc_W"] + P["enc_b"], 0.0)
    cache = {"F": F, "s": s, "H": [H], "Z": [], "M": [], "U": [], "P": P}
    for l in range(layers):
        Z = rnd(H) @ rnd(P["Wab%d" % l])
        pre = Z[g.ei, :HID] + Z[g.ej, HID:] + g.e[:, None] * P["we%d" % l] + P["bm%d" % l]
        M = g.S @ np.maximum(pre, 0.0)
        U = np.maximum(rnd(np.hstack([H, M])) @ rnd(P["Wu%d" % l]) + P["bu%d" % l], 0.0)
        H = H + U
        cache["Z"].append(Z)
        cache["M"].append(M)
        cache["U"].append(U)
        cache["H"].append(H)
    d = s * (H @ P["dec_w"] + P["dec_b"])
    xb = base(x_n, x_p, ext)
    r = b - g.A @ xb
    Ad = g.A @ d
    dAd = float(d @ Ad)
    if obj == 1:
        q = float(Ad @ Ad)
        alpha = float(Ad @ r) / q if q > 0 else 0.0
    else:
        alpha = float(d @ r) / dAd if dAd > 0 else 0.0
    cache.update(d=d, Ad=Ad, dAd=dAd, xb=xb, r=r)
    return xb + alpha * d, cache


def loss_residual(g, cache):
    """Minimum-residual objective: L and dL/dd (see the module text)."""
    Ad, r = cache["Ad"], cache["r"]
    R = float(r @ r)
    q = float(Ad @ Ad)
    if not R > 0:  # the base solves the system
        return 0.0, np.zeros_like(Ad)
    if not q > 0:  # no direction: nothing removed
        return 1.0, np.zeros_like(Ad)
    c = float(Ad @ r)
    L = 1.0 - c * c / (q * R)
    w = (-2.0 * c / (q * R)) * r + (2.0 * c * c / (q * q * R)) * Ad
    return L, g.A @ w


def loss(g, cache, x_star):
    """L and dL/dd of the line-searched direction (see the module text)."""
    d, Ad, dAd = cache["d"], cache["Ad"], cache["dAd"]
    eb = x_star - cache["xb"]
    gv = g.A @ eb
    E0 = float(eb @ gv)
    if not E0 > 0:  # the base is already the solution
        return 0.0, np.zeros_like(d)
    if not dAd > 0:  # no direction: nothing removed
        return 1.0, np.zeros_like(d)
    dg = float(d @ gv)
    L = 1.0 - dg * dg / (dAd * E0)
    return L, -2.0 * dg / (dAd * E0) * (gv - (dg / dAd) * Ad)


def backward(g, cache, layers, gd, bf=False):
    """Gradient of the flat parameter vector given dL/dd = gd."""
    P = cache["P"]
    rnd = bf16 if bf else (lambda a: a)
    s = cache["s"]
    G = {}
    dd = s * gd
    Hl = cache["H"][-1]
    G["dec_w"] = Hl.T @ dd
    G["dec_b"] = dd.sum()
    dH = np.outer(dd, P["dec_w"])
    for l in reversed(range(layers)):
        H, Z, M, U = cache["H"][l], cache["Z"][l], cache["M"][l], cache["U"][l]
        gu = dH * (U > 0)
        HM = np.hstack([H, M])
        G["Wu%d" % l] = rnd(HM).T @ rnd(gu)
        G["bu%d" % l] = gu.sum(0)
        T = rnd(gu) @ rnd(P["Wu%d" % l]).T
        dM = T[:, HID:]
        dHm = dH + T[:, :HID]
        pre = Z[g.ei, :HID] + Z[g.ej, HID:] + g.e[:, None] * P["we%d" % l] + P["bm%d" % l]
        sij = dM[g.ei] * (pre > 0)
        dZ = np.hstack([g.S @ sij, g.St @ sij])
        G["we%d" % l] = (sij * g.e[:, None]).sum(0)
        G["bm%d" % l] = sij.sum(0)
        G["Wab%d" % l] = rnd(H).T @ rnd(dZ)
        dH = dHm + rnd(dZ) @ rnd(P["Wab%d" % l]).T
    g0 = dH * (cache["H"][0] > 0)
    G["enc_W"] = cache["F"].T @ g0
    G["enc_b"] = g0.sum(0)
    return pack(G, layers)


def loss_and_grad(g, theta, layers, x_n, x_p, b, x_star, bf=False, ext=0, obj=0):
    x0, cache = forward(g, theta, layers, x_n, x_p, b, bf, ext, obj)
    L, gd = loss_residual(g, cache) if obj == 1 else loss(g, cache, x_star)
    return L, backward(g, cache, layers, gd, bf), x0


def init_params(layers, seed=0, h=HID):
    """A random (Glorot-uniform) parameter vector for tests."""
    rng = np.random.default_rng(seed)
    P = {"enc_W": rng.uniform(-1, 1, (FIN, h)) * np.sqrt(6.0 / (FIN + h)),
         "enc_b": rng.uniform(-0.1, 0.1, h), "dec_w": rng.uniform(-1, 1, h) * np.sqrt(6.0 / (h + 1)),
         "dec_b": 0.01}
    for l in range(layers):
        P["Wab%d" % l] = rng.uniform(-1, 1, (h, 2 * h)) * np.sqrt(6.0 / (3 * h))
        P["we%d" % l] = rng.uniform(-1, 1, h)
        P["bm%d" % l] = rng.uniform(-0.1, 0.1, h)
        P["Wu%d" % l] = rng.uniform(-1, 1, (2 * h, h)) * np.sqrt(6.0 / (3 * h))
        P["bu%d" % l] = rng.uniform(-0.1, 0.1, h)
    return pack(P, layers)
