"""Dense FP64 torch assembly of the 3D Scharfetter-Gummel / upwind-diffusion
operator (rows of mps1d/source/convdiff.f90:287-293 in each direction, BC
rows of mgridlib/discretepde.f90:810-863), differentiable by torch autograd.
The independent reference the adjoint and discharge tests solve against."""
import torch


def bernoulli_t(x):
    """B(x) = x / (e^x - 1), the series below |x| < 1e-3, and for x > 0 the
    form x e^-x / (1 - e^-x), whose value and derivative stay finite at any
    Peclet number (x / expm1(x) differentiates to inf / inf above x ~ 709).
    Each branch sees only inputs from its own range, so the branches torch
    does not select contribute zero gradients rather than NaN."""
    small = x.abs() < 1e-3
    pos = x > 0
    xp = torch.where(pos & ~small, x, torch.ones_like(x))
    xn = torch.where(~pos & ~small, x, -torch.ones_like(x))
    bp = xp * torch.exp(-xp) / (-torch.expm1(-xp))
    bn = xn / torch.expm1(xn)
    ser = 1.0 - 0.5 * x + x * x * (1.0 / 12.0 - x * x / 720.0)
    return torch.where(small, ser, torch.where(pos, bp, bn))


def dense_sg(n, h, D, vel, codes, dtinv, mean="harmonic"):
    """Dense FP64 A(v) of the 3D SG operator (interior (nz,ny,nx) tensors,
    row-major cell index i + nx*(j + ny*k)), built with torch ops so that
    autograd differentiates it."""
    nx, ny, nz = n
    N = nx * ny * nz
    A = torch.zeros((N, N), dtype=torch.float64, device=D.device)
    idx = torch.arange(N, device=D.device).reshape(nz, ny, nx)
    diag = torch.zeros((nz, ny, nx), dtype=torch.float64, device=D.device) + dtinv
    for axis, v, lo, hi in ((2, vel[0], codes[0], codes[1]), (1, vel[1], codes[2], codes[3]),
                            (0, vel[2], codes[4], codes[5])):
        sl = [slice(None)] * 3
        sr = [slice(None)] * 3
        sl[axis] = slice(0, -1)
        sr[axis] = slice(1, None)
        sl, sr = tuple(sl), tuple(sr)
        cL, cR, DL, DR = v[sl], v[sr], D[sl], D[sr]
        ch = 0.5 * (cL + cR)
        Dh = 2.0 * DL * DR / (DL + DR + 1e-10) if mean == "harmonic" else 0.5 * (DL + DR)
        Pe = ch * h / Dh
        g = Dh / h
        wL = g * bernoulli_t(-Pe)
        wR = -g * bernoulli_t(Pe)
        iL, iR = idx[sl].reshape(-1), idx[sr].reshape(-1)
        # row L: + (wL uL + wR uR)/h ; row R: - (...)/h
        A = A.index_put((iL, iR), (wR / h).reshape(-1), accumulate=True)
        A = A.index_put((iR, iL), (-wL / h).reshape(-1), accumulate=True)
        dL = torch.zeros_like(diag)
        dL[sl] = dL[sl] + wL / h
        dL[sr] = dL[sr] - wR / h
        diag = diag + dL
        # boundary rows: DIRC = 2D/h^2 + upwind outflow, ZGRD = outflow
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[axis] = slice(0, 1)
        last[axis] = slice(-1, None)
        first, last = tuple(first), tuple(last)
        bd = torch.zeros_like(diag)
        vb = v[first]
        if lo == "DIRC":
            bd[first] = 2 * D[first] / h ** 2 + torch.where(vb < 0, -vb / h, torch.zeros_like(vb))
        elif lo == "ZGRD":
            bd[first] = -vb / h
        vb = v[last]
        if hi == "DIRC":
            bd[last] = bd[last] + 2 * D[last] / h ** 2 + torch.where(vb > 0, vb / h, torch.zeros_like(vb))
        elif hi == "ZGRD":
            bd[last] = bd[last] + vb / h
        diag = diag + bd
    A = A + torch.diag(diag.reshape(-1))
    return A
