"""CPU oracle of the learned preconditioner's network (test infrastructure:
only tests, smoke() and bench.py's CPU leg use it).

The network has no counterpart in the reference; this is a plain numpy FP64
statement of the forward pass and its hand-derived backward pass, with the
same parameter layout as the GPU kernels:

    s      = rms_i |r_i|,  x_in_i = [r_i/s, n_i, p_i]
    h^0    = relu(W_in x_in + b_in)
    layer: a = W_a h;  m_i = mean_{j in N(i)} relu(a_j + Bd dl_ij + c);
           g_S = mean_{i in S} h_i;  h <- h + relu(W_u [h, m, g_S(i)] + d)
    z      = r + s (W_out h^L + b_out)
    loss   = mean_b |z_b - v_b|^2 / |v_b|^2

With bf16=True every matrix-product operand is rounded to BF16 (round to
nearest even, from its FP32 value) as the GPU's matrix-core products do, so
the GPU results can be compared closely; with bf16=False the backward pass is
the exact gradient of the FP64 forward pass, which the tests pin by central
finite differences.
"""
import numpy as np

W = 64
INPAD = 16
LAYERS = 3
UW = 3 * W
OFF_WIN = 0
OFF_BIN = OFF_WIN + W * INPAD
OFF_L0 = OFF_BIN + W
L_WA = 0
L_BD = L_WA + W * W
L_CC = L_BD + W * 4
L_WU = L_CC + W
L_D = L_WU + W * UW
L_SIZE = L_D + W
OFF_WOUT = OFF_L0 + LAYERS * L_SIZE
OFF_BOUT = OFF_WOUT + 3 * W
NPARAMS = OFF_BOUT + 4


def bf16(x):
    """FP32 value of x rounded to BF16, round to nearest even (finite x)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed, scale_out=0.0):
    """He-style random weights; the output layer starts at scale_out (0: M = I)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    p[OFF_WIN:OFF_WIN + W * INPAD] = (rng.standard_normal((W, INPAD)) * np.sqrt(2.0 / 9)
                                     * (np.arange(INPAD) < 9)).ravel()
    for l in range(LAYERS):
        b = OFF_L0 + l * L_SIZE
        p[b + L_WA:b + L_WA + W * W] = rng.standard_normal(W * W) * np.sqrt(1.0 / W)
        bd = rng.standard_normal((W, 4)) * 0.3
        bd[:, 3] = 0.0
        p[b + L_BD:b + L_BD + W * 4] = bd.ravel()
        p[b + L_WU:b + L_WU + W * UW] = rng.standard_normal(W * UW) * np.sqrt(1.0 / UW)
    p[OFF_WOUT:OFF_WOUT + 3 * W] = rng.standard_normal(3 * W) * scale_out / np.sqrt(W)
    return p


class Graph:
    """Point set in the solver's unknown order and its k-nearest-neighbour
    graph nbr (n x k, from the library), with the edge offsets dl_ij =
    (x_j - x_i)/rho_i and the geometric features of the GPU (FP32 values)."""

    def __init__(self, n1, n2, x, nrml, ctr, size, nbr):
        self.n1, self.n2 = n1, n2
        self.npts = n1 + n2
        x = np.asarray(x, np.float64).reshape(-1, 3)
        self.nbr = np.asarray(nbr, np.int64).reshape(self.npts, -1)
        self.k = self.nbr.shape[1]
        d = x[self.nbr] - x[:, None, :]
        rho = np.linalg.norm(d, axis=2).mean(axis=1)
        rho[rho <= 0] = 1.0
        self.dl = (d / rho[:, None, None]).astype(np.float32).astype(np.float64)
        body = (np.arange(self.npts) >= n1).astype(int)
        c = np.asarray(ctr, np.float64).reshape(2, 3)[body]
        sz = np.asarray(size, np.float64)[body]
        self.geo = np.concatenate([np.asarray(nrml, np.float64).reshape(-1, 3),
                                   (x - c) / sz[:, None]], axis=1).astype(np.float32).astype(np.float64)
        self.seg = body


class Net:
    def __init__(self, params, bf=True):
        self.p = np.asarray(params, np.float32).astype(np.float64)
        self.bf = bf

    def _r(self, a):
        return bf16(a) if self.bf else a

    def _mm(self, a, b):
        return self._r(a) @ self._r(b)

    def _layer(self, l):
        b = OFF_L0 + l * L_SIZE
        p = self.p
        return (p[b + L_WA:b + L_WA + W * W].reshape(W, W), p[b + L_BD:b + L_BD + W * 4].reshape(W, 4)[:, :3],
                p[b + L_CC:b + L_CC + W], p[b + L_WU:b + L_WU + W * UW].reshape(W, UW), p[b + L_D:b + L_D + W])

    def forward(self, G, r):
        """r: (batch, 3*npts).  Returns z and the cache for backward."""
        p = self.p
        r = np.asarray(r, np.float64).reshape(-1, G.npts, 3)
        B = r.shape[0]
        s = np.sqrt((r ** 2).sum(axis=(1, 2)) / G.npts)
        inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
        xin = np.zeros((B, G.npts, INPAD))
        xin[:, :, :3] = (r * inv[:, None, None]).astype(np.float32)
        xin[:, :, 3:9] = G.geo[None]
        Win = p[OFF_WIN:OFF_WIN + W * INPAD].reshape(W, INPAD)
        pre0 = self._mm(xin, Win.T) + p[OFF_BIN:OFF_BIN + W]
        h = np.maximum(pre0, 0.0)
        cache = dict(B=B, r=r, s=s, xin=xin, pre0=pre0, layers=[])
        for l in range(LAYERS):
            Wa, Bd, cc, Wu, d = self._layer(l)
            a = self._mm(h, Wa.T)
            g = np.stack([h[:, G.seg == 0].mean(axis=1), h[:, G.seg == 1].mean(axis=1)], axis=1)  # (B,2,W)
            pe = a[:, G.nbr, :] + (G.dl @ Bd.T)[None] + cc           # (B, n, k, W)
            m = np.maximum(pe, 0.0).mean(axis=2)
            U = np.concatenate([h, m, g[:, G.seg, :]], axis=2)
            pre = self._mm(U, Wu.T) + d
            cache["layers"].append(dict(h=h, a=a, pe=pe, U=U, pre=pre))
            h = h + np.maximum(pre, 0.0)
        Wout = p[OFF_WOUT:OFF_WOUT + 3 * W].reshape(3, W)
        o = self._mm(h, Wout.T)
        z = r + s[:, None, None] * (o + p[OFF_BOUT:OFF_BOUT + 3])
        cache["hL"] = h
        return z.reshape(B, -1), cache

    def loss_grad(self, G, r, v):
        z, c = self.forward(G, r)
        B = c["B"]
        v = np.asarray(v, np.float64).reshape(B, -1)
        vn2 = (v ** 2).sum(axis=1)
        w = np.where(vn2 > 0, 1.0 / (np.where(vn2 > 0, vn2, 1.0) * B), 0.0)
        diff = z - v
        loss = float(((diff ** 2).sum(axis=1) * w).sum())
        p = self.p
        g = np.zeros(NPARAMS)
        dO = (2.0 * diff * w[:, None] * c["s"][:, None]).reshape(B, G.npts, 3)
        g[OFF_BOUT:OFF_BOUT + 3] = dO.sum(axis=(0, 1))
        hL = c["hL"]
        Wout = p[OFF_WOUT:OFF_WOUT + 3 * W].reshape(3, W)
        g[OFF_WOUT:OFF_WOUT + 3 * W] = np.einsum("bni,bnj->ij", self._r(dO), self._r(hL)).ravel()
        dh = self._mm(dO, Wout)
        nseg = np.array([(G.seg == 0).sum(), (G.seg == 1).sum()])
        for l in reversed(range(LAYERS)):
            Wa, Bd, cc, Wu, d = self._layer(l)
            L = c["layers"][l]
            gb = OFF_L0 + l * L_SIZE
            dpre = dh * (L["pre"] > 0)
            g[gb + L_D:gb + L_D + W] = dpre.sum(axis=(0, 1))
            g[gb + L_WU:gb + L_WU + W * UW] = np.einsum("bni,bnj->ij", self._r(dpre), self._r(L["U"])).ravel()
            dU = self._mm(dpre, Wu)
            dg = np.stack([dU[:, G.seg == 0, 2 * W:].sum(axis=1), dU[:, G.seg == 1, 2 * W:].sum(axis=1)], axis=1)
            dh = dh + dU[:, :, :W] + dg[:, G.seg, :] / nseg[G.seg][None, :, None]
            dm = dU[:, :, W:2 * W]
            de = (L["pe"] > 0) * (dm[:, :, None, :] / G.k)          # (B, n, k, W)
            dA = np.zeros_like(L["a"])
            for b in range(B):
                np.add.at(dA[b], G.nbr.ravel(), de[b].reshape(-1, W))
            gBd = np.einsum("bnkc,nkd->cd", de, G.dl)
            bd = np.zeros((W, 4))
            bd[:, :3] = gBd
            g[gb + L_BD:gb + L_BD + W * 4] = bd.ravel()
            g[gb + L_CC:gb + L_CC + W] = de.sum(axis=(0, 1, 2))
            g[gb + L_WA:gb + L_WA + W * W] = np.einsum("bni,bnj->ij", self._r(dA), self._r(L["h"])).ravel()
            dh = dh + self._mm(dA, Wa)
        dpre0 = dh * (c["pre0"] > 0)
        g[OFF_BIN:OFF_BIN + W] = dpre0.sum(axis=(0, 1))
        g[OFF_WIN:OFF_WIN + W * INPAD] = np.einsum("bni,bnj->ij", self._r(dpre0), self._r(c["xin"])).ravel()
        return loss, g, z
