"""CPU restatement of the learned current denoiser (test infrastructure).

The network of tristan-mp-pu-master_densdecomp_amd/csrc/denoise.hip in
float64 numpy: periodic 3x3 (2D) or 3x3x3 (3D) convolutions, GELU (tanh
form), a vector potential A on the B grid and the correction curl_b(A) with
the backward-difference curl of the E update (code/fields.F90:836-838).
`bf16=True` rounds to bfloat16 at the points where the GPU stores BF16
(input, weights, hidden activations, backward deltas), so GPU and oracle
differ only by FP32 accumulation order.  This module has no counterpart in
the reference (the denoiser is new); it is pinned by finite differences of
its own analytic gradient (tests/test_denoiser_oracle.py) and by the
identity div_b(curl_b A) = 0.
"""
import numpy as np


def bf16(a):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    x = np.ascontiguousarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def taps(dim):
    """(dx, dy, dz) per tap, x fastest, as the GPU's neighbour()"""
    out = []
    n = 27 if dim == 3 else 9
    for t in range(n):
        out.append((t % 3 - 1, (t // 3) % 3 - 1, (t // 9 - 1) if dim == 3 else 0))
    return out


def shift(a, d):
    """a[z+dz, y+dy, x+dx] periodically; a has shape (nz, ny, nx, C)"""
    dx, dy, dz = d
    return np.roll(a, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def unshift(a, d):
    dx, dy, dz = d
    return np.roll(a, shift=(dz, dy, dx), axis=(0, 1, 2))


def gelu(z):
    k0, k1 = 0.7978845608, 0.044715
    return 0.5 * z * (1.0 + np.tanh(k0 * (z + k1 * z ** 3)))


def gelu_grad(z):
    k0, k1 = 0.7978845608, 0.044715
    u = k0 * (z + k1 * z ** 3)
    th = np.tanh(u)
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * k0 * (1.0 + 3.0 * k1 * z * z)


def curl_b(A, dim):
    """A (nz, ny, nx, 3) -> curl with backward differences"""
    def dm(f, axis):  # f - f shifted by -1 along axis (backward difference)
        return f - np.roll(f, 1, axis=axis)
    Ax, Ay, Az = A[..., 0], A[..., 1], A[..., 2]
    cx = dm(Az, 1)
    cy = -dm(Az, 2)
    cz = dm(Ay, 2) - dm(Ax, 1)
    if dim == 3:
        cx = cx - dm(Ay, 0)
        cy = cy + dm(Ax, 0)
    return np.stack([cx, cy, cz], axis=-1)


def curl_b_adj(g, dim):
    def dp(f, axis):  # f - f shifted by +1 (adjoint of the backward difference)
        return f - np.roll(f, -1, axis=axis)
    gx, gy, gz = g[..., 0], g[..., 1], g[..., 2]
    ax = -dp(gz, 1)
    ay = dp(gz, 2)
    az = dp(gx, 1) - dp(gy, 2)
    if dim == 3:
        ay = ay - dp(gx, 0)
        ax = ax + dp(gy, 0)
    return np.stack([ax, ay, az], axis=-1)


class Net:
    """layers: list of (W (cout, taps, cin), b (cout,)); first cin = 3"""

    def __init__(self, dim, params=None, width=None, nlayers=None):
        self.dim = dim
        self.width = width or (32 if dim == 2 else 16)
        self.nl = nlayers or (4 if dim == 2 else 3)
        self.nt = 27 if dim == 3 else 9
        self.shapes = []
        cin = 3
        for l in range(self.nl):
            cout = 3 if l == self.nl - 1 else self.width
            self.shapes.append((cout, self.nt, cin))
            cin = cout
        if params is not None:
            self.set_flat(params)

    def nparams(self):
        return sum(c * t * i + c for c, t, i in self.shapes)

    def set_flat(self, flat):
        flat = np.asarray(flat, dtype=np.float64)
        self.layers = []
        o = 0
        for c, t, i in self.shapes:
            W = flat[o:o + c * t * i].reshape(c, t, i)
            o += c * t * i
            b = flat[o:o + c]
            o += c
            self.layers.append((W, b))
        assert o == flat.size

    def flat_grads(self, grads):
        return np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])

    def conv(self, x, W, b, r):
        out = np.zeros(x.shape[:3] + (W.shape[0],))
        Wr = r(W)
        for t, d in enumerate(taps(self.dim)):
            out += shift(x, d) @ Wr[:, t, :].T
        return out + (b if b is not None else 0.0)

    def conv_T(self, dz, W, r):
        out = np.zeros(dz.shape[:3] + (W.shape[2],))
        Wr = r(W)
        for t, d in enumerate(taps(self.dim)):
            out += unshift(dz @ Wr[:, t, :], d)
        return out

    def loss_and_grad(self, jlo, jt, s, bf=True, grad=True, lam=0.0):
        """jlo, jt: (3, nz, ny, nx) current arrays; returns loss, flat grad.
        loss = mean(d^2) + lam <d,t>^2 / (N <t,t>), d = prediction - target,
        t = target (both / s): the second term penalises the error along the
        target, i.e. a shrunk or inflated amplitude."""
        r = bf16 if bf else _id
        f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf else _id
        inv_s = np.float32(1.0) / np.float32(s) if bf else 1.0 / s
        xl = np.moveaxis(f32(np.asarray(jlo, np.float64) * inv_s), 0, -1)
        xt = np.moveaxis(f32(np.asarray(jt, np.float64) * inv_s), 0, -1)
        h = r(xl)
        zs, hs = [], [h]
        for l, (W, b) in enumerate(self.layers):
            z = self.conv(hs[-1], W, b, r)
            zs.append(z)
            if l < self.nl - 1:
                hs.append(r(gelu(z)))
        A = zs[-1]
        c = curl_b(A, self.dim)
        d = xl + c - xt
        N = d.size
        tt = float(np.sum(xt * xt))
        dt = float(np.sum(d * xt))
        bias = dt / tt if (lam and tt > 0.0) else 0.0
        loss = float(np.sum(d * d) / N) + lam * bias * dt / N
        if not grad:
            return loss, None, xl + c
        g = 2.0 / N * (d + lam * bias * xt)
        dz = r(curl_b_adj(g, self.dim))
        grads = [None] * self.nl
        for l in range(self.nl - 1, -1, -1):
            W, b = self.layers[l]
            x = hs[l]
            gW = np.zeros_like(W)
            for t, dd in enumerate(taps(self.dim)):
                xs = shift(x, dd).reshape(-1, x.shape[-1])
                gW[:, t, :] = dz.reshape(-1, dz.shape[-1]).T @ xs
            gb = dz.reshape(-1, dz.shape[-1]).sum(0)
            grads[l] = (gW, gb)
            if l > 0:
                dh = self.conv_T(dz, W, r)
                dz = r(dh * gelu_grad(zs[l - 1]))
        return loss, self.flat_grads(grads), xl + c
