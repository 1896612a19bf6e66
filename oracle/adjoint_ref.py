# SPDX-License-Identifier: Apache-2.0
"""The reference of the adjoint: the decomposition's sift map with its decisions frozen, restated in
PyTorch float64 on the CPU, so that autograd gives the FP64 reference gradient.

Test infrastructure only (like the rest of oracle/): the tests import it as the checker; the
product package never does.

The frozen map replays one element's decomposition from a tape (oracle.decompose_tape, or the
library's own): per mode the windows w and W, the rung ladder with its used / segment-point bits,
the scale field L, the lower rung lo of every voxel, the tie tolerance, which sifts ran and which
were accepted, and each sift's amplitude floor thr. Everything else is recomputed from the input as
a function of it: the orientation n, c (Riesz transform of the residual in the mode's band, the
structure tensor, its top eigenvector; frozen with orientation=False), the element d, e, o = e n,
the hat weight t of the upper rung, the rung cubes (max / min / box with half-sample reflection),
the 5-point gathers with D-linear interpolation, the continuation weight, G1's ramp and blend, and
the sifts. The steps follow cpp/src/method/driver.cpp:540-689 (element.cu:166-292,
element_op.cuh:86-282, sift.cu:55-232, orientation.cu:86-404), as the oracle's decompose_one
restates them. Max and min route their gradient to the position they select (the first in the
window where values tie); clip, max(., 0) and the ramps pass it where their argument lies inside.
"""

from __future__ import annotations

import math

import numpy as np
import torch

Q, SPAN, T0, T1, RAMP_K0, RAMP_K1, U_OPS, S_OPS = 4.0, 1.0, 0.2, 0.5, 0.3, 0.6, 2, 1


def reflect_any(i: torch.Tensor, n: int) -> torch.Tensor:
    m = torch.remainder(i, 2 * n)
    return torch.where(m < n, m, 2 * n - 1 - m)


# ---- the cube (cpp/src/method/common.cpp:65-136) -------------------------------------------------
def _axis_window(f: torch.Tensor, a: int, w: int, op: str) -> torch.Tensor:
    n = f.shape[a]
    r = min((w - 1) // 2, n - 1)
    if r == 0:
        return f
    idx = reflect_any(torch.arange(-r, n + r), n)
    ext = torch.index_select(f, a, idx)
    win = ext.unfold(a, 2 * r + 1, 1)  # (..., n, ...) with the window last
    if op == "max":
        return win.max(dim=-1).values
    if op == "min":
        return win.min(dim=-1).values
    return win.sum(dim=-1)


def cube(f: torch.Tensor, op: str, w: int) -> torch.Tensor:
    D = f.dim()
    if op == "box":
        prod = 1.0
        for a in range(D):
            prod *= 2 * min((w - 1) // 2, f.shape[a] - 1) + 1
        out = f
        for a in range(D):
            out = _axis_window(out, a, w, "sum")
        return out * (1.0 / prod)
    out = f
    for a in range(D - 1, -1, -1):
        out = _axis_window(out, a, w, op)
    return out


# ---- D-linear interpolation of the half-sample reflected periodic extension --------------------
def interp(C: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """C (spatial), y (D, *spatial) coordinates: the interpolated values (element.cu interp)."""
    D = C.dim()
    fl = torch.floor(y.detach())
    w1 = y - fl
    i0 = fl.to(torch.int64)
    flat = C.reshape(-1)
    st = [math.prod(C.shape[a + 1:]) for a in range(D)]
    acc = torch.zeros_like(y[0])
    for corner in range(1 << D):
        off = torch.zeros_like(i0[0])
        wt = torch.ones_like(y[0])
        for a in range(D):
            hi = (corner >> (D - 1 - a)) & 1
            idx = reflect_any(i0[a] + hi, C.shape[a])
            off = off + idx * st[a]
            wt = wt * (w1[a] if hi else 1.0 - w1[a])
        acc = acc + flat[off.reshape(-1)].reshape(off.shape) * wt
    return acc


# ---- the element and its operators ---------------------------------------------------------------
class Element:
    def __init__(self, mode: dict, L, c, n, D: int):
        self.rung_w = mode["rung_w"]
        self.K = len(self.rung_w)
        self.used = mode["used"]
        self.pts = mode["pts"]
        self.lo = torch.from_numpy(mode["lo"].astype(np.int64))
        Lt = torch.as_tensor(L, dtype=torch.float64)
        f = (1.0 - c) + c / Q
        d = Lt * f
        ln = Lt * (1.0 + (SPAN - 1.0) * c)
        e = 0.5 * torch.clamp(ln - d, min=0.0)
        self.o = e[None] * n if D >= 2 else None
        if self.K > 1:
            lw = torch.log(torch.tensor(self.rung_w, dtype=torch.float64))
            dc = torch.clamp(d, min=float(self.rung_w[0]), max=float(self.rung_w[-1]))
            lg = torch.log(dc)
            self.t = (lg - lw[self.lo]) / (lw[self.lo + 1] - lw[self.lo])
        else:
            self.t = torch.zeros_like(d)
        self.D = D
        grids = torch.meshgrid(*[torch.arange(s, dtype=torch.float64) for s in L.shape], indexing="ij")
        self.x = torch.stack(grids)

    def op(self, f: torch.Tensor, kind: str) -> torch.Tensor:
        """kind 'max' (Dil), 'min' (Ero) or 'box' (Avg)."""
        out = torch.zeros_like(f)
        wl = 1.0 - self.t
        for j in range(self.K):
            if not (self.used >> j) & 1:
                continue
            C = cube(f, kind, self.rung_w[j])
            acc = C
            if self.D >= 2 and (self.pts >> j) & 1:
                for k in (1, 2):
                    for sg in (1.0, -1.0):
                        v = interp(C, self.x + sg * (k / 2.0) * self.o)
                        if kind == "max":
                            acc = torch.maximum(acc, v)
                        elif kind == "min":
                            acc = torch.minimum(acc, v)
                        else:
                            acc = acc + v
                if kind == "box":
                    acc = acc / 5.0
            wgt = torch.where((self.lo == j) & (wl > 0), wl, torch.zeros_like(wl))
            wgt = wgt + torch.where((self.lo + 1 == j) & (self.t > 0), self.t, torch.zeros_like(wl))
            out = out + wgt * acc
        return out

    def envelope_mean(self, h):
        u = self.op(h, "max")
        lo = self.op(h, "min")
        return self.op(0.5 * (u + lo), "box"), u, lo


# ---- the orientation (orientation.cu:86-404) -----------------------------------------------------
def _dct_mats(n: int):
    k = torch.arange(n, dtype=torch.float64)[:, None]
    x = torch.arange(n, dtype=torch.float64)[None, :]
    ck = torch.full_like(k, math.sqrt(2.0 / n))
    ck[0] = math.sqrt(1.0 / n)
    ang = math.pi * k * (2 * x + 1) / (2 * n)
    return ck * torch.cos(ang), ck * torch.sin(ang)  # C[k, x], S[k, x]


def _apply_axis(f, M, a):
    return torch.movedim(torch.tensordot(M, torch.movedim(f, a, 0), dims=1), 0, a)


def band_two_sigma2() -> float:
    l2 = math.log(2.0)
    sg = l2 / (2.0 * math.sqrt(2.0 * l2))
    return 2.0 * (sg * sg)


def orientation(r: torch.Tensor, w: int):
    D = r.dim()
    shape = r.shape
    mats = [_dct_mats(s) for s in shape]
    Y = r - r.mean()
    X = Y
    for a in range(D):
        X = _apply_axis(X, mats[a][0], a)
    grids = torch.meshgrid(*[torch.arange(s, dtype=torch.float64) for s in shape], indexing="ij")
    u = [grids[a] * (1.0 / (2.0 * shape[a])) for a in range(D)]
    rad = torch.sqrt(sum(ua * ua for ua in u))
    safe = torch.where(rad > 0, rad, torch.ones_like(rad))
    l = torch.log(safe * w)
    G = torch.where(rad > 0, torch.exp(-(l * l) / band_two_sigma2()), torch.zeros_like(rad))
    Xb = X * G
    ua = [torch.where(rad > 0, u[a] / safe, torch.zeros_like(rad)) for a in range(D)]

    def inverse(m, odd):
        y = m * Xb
        for a in range(D):
            M = mats[a][1] if (odd >> a) & 1 else mats[a][0]
            y = _apply_axis(y, M.T, a)
        return y

    R1 = [inverse(ua[a], 1 << a) for a in range(D)]
    R2 = {}
    for a in range(D):
        for b in range(a, D):
            R2[(a, b)] = inverse(-ua[a] * ua[a] if a == b else ua[a] * ua[b], 0 if a == b else (1 << a) | (1 << b))
            R2[(b, a)] = R2[(a, b)]
    Tm = torch.empty(shape + (D, D), dtype=torch.float64)
    rows = []
    for a in range(D):
        row = []
        for c in range(D):
            v = R1[a] * R1[c]
            for k in range(D):
                v = v + R2[(a, k)] * R2[(k, c)]
            row.append(v)
        rows.append(torch.stack(row, dim=-1))
    Tm = torch.stack(rows, dim=-2)
    evals, evecs = torch.linalg.eigh(Tm)
    mu1, mu2 = evals[..., -1], evals[..., -2]
    n = torch.movedim(evecs[..., :, -1], -1, 0)
    ssum = mu1 + mu2
    cc = torch.where(ssum > 0, (mu1 - mu2) / torch.where(ssum > 0, ssum, torch.ones_like(ssum)),
                     torch.zeros_like(ssum))
    return n, torch.clamp(cc, 0.0, 1.0)


# ---- the frozen decomposition ------------------------------------------------------------------
def frozen_decompose(x: torch.Tensor, tape: dict, *, orientation_grad: bool = True, tiny: float = 1e-300):
    """The rows (P, *spatial) of x (float64, spatial shape) under the tape's frozen decisions."""
    D = x.dim()
    kexp = tape["kexp"]
    r = x * (2.0 ** -kexp)
    rows = []
    fl_k0, kden = RAMP_K0, max(RAMP_K1 - RAMP_K0, 1e-12)
    den_t = max(T1 - T0, 1e-12)
    for mode in tape["modes"]:
        if D >= 2:
            n, c = orientation(r, mode["w"])
            if not orientation_grad:
                n, c = n.detach(), c.detach()
        else:
            n, c = None, torch.zeros_like(r)
        el = Element(mode, torch.from_numpy(np.asarray(mode["L"], np.float64)), c, n, D)
        tol = mode["tol"]
        W = mode["W"]
        m1, u1, l1 = el.envelope_mean(r)
        h = r - m1
        a = None
        if tape["sifts"] > 1:
            R = torch.clamp(u1 - l1, min=0.0)
            V = R
            for _ in range(U_OPS):
                V = el.op(V, "max")
            safe = torch.where(V > tol, V, torch.ones_like(V))
            rho = torch.where(V > tol, torch.clamp(R / safe, 0.0, 1.0), torch.zeros_like(V))
            gw = torch.clamp((rho - T0) / den_t, 0.0, 1.0)
            for _ in range(S_OPS):
                gw = el.op(gw, "box")
            a = torch.clamp(gw, 0.0, 1.0)
        for j in range(mode["ran"]):
            plain, up, lo = el.envelope_mean(h)
            e0 = el.op(0.5 * (up - lo), "box")
            emax, emin = cube(e0, "max", W), cube(e0, "min", W)
            fl = max(tol, tiny)
            ssum = emax + emin
            den = torch.where(ssum >= fl, ssum, torch.full_like(ssum, fl))
            kap = (emax - emin) / den
            kr = torch.clamp((kap - fl_k0) / kden, 0.0, 1.0)
            gk = torch.clamp(cube(kr, "box", W), 0.0, 1.0)
            e = cube(e0, "box", W)
            thr = mode["thr"][j]
            e = torch.where(e >= thr, e, torch.full_like(e, thr))
            q = h / e
            Mn, _, _ = el.envelope_mean(q)
            mA = torch.minimum(torch.maximum(e * Mn, lo), up)
            blend = gk * mA + (1.0 - gk) * plain
            hn = h - a * blend
            if mode["accepted"][j]:
                h = hn
        rows.append(h * (2.0 ** kexp))
        r = r - h
    rows.append(r * (2.0 ** kexp))
    return torch.stack(rows)


def loss_weights(P: int, seed: int, shape) -> tuple[torch.Tensor, torch.Tensor]:
    """The differentiation workload's loss: sum_k w_k ||row_k - target_k||^2 with w_k ~ U(0.5, 2) and
    targets ~ N(0, 1) from default_rng(seed); rows k < P (the residue row included)."""
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.uniform(0.5, 2.0, P))
    tgt = torch.from_numpy(rng.standard_normal((P,) + tuple(shape)))
    return w, tgt


def reference_gradient(x: np.ndarray, tape: dict, seed: int, *, orientation_grad: bool = True):
    """(loss, d loss / d x) in float64 of the frozen map at x under the tape."""
    xt = torch.from_numpy(np.asarray(x, np.float64)).clone().requires_grad_(True)
    rows = frozen_decompose(xt, tape, orientation_grad=orientation_grad)
    w, tgt = loss_weights(rows.shape[0], seed, x.shape)
    loss = (w.reshape((-1,) + (1,) * x.ndim) * (rows - tgt) ** 2).sum()
    loss.backward()
    return float(loss.detach()), xt.grad.numpy().copy(), rows.detach().numpy()
