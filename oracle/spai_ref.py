"""FP64 CPU oracle of the learned SPAI preconditioner (network, preconditioner, probe loss).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP forward pass and the hand-written
backward pass.  Same definitions as lagrangianvoronoi.jl_amd/csrc/lv_spai.hip, evaluated in
float64 with torch on the CPU; gradients come from autograd and are themselves checked against
central finite differences.  No counterpart exists in the reference.
"""
import numpy as np
import torch

H = 64
R_FLOOR = 1e-4
OCLAMP = 8.0
OFF_WN, OFF_BN, OFF_WE, OFF_BE = 0, 256, 320, 576


def OFF_W(l):
    return 640 + l * (192 * 64 + 64)


def OFF_B(l):
    return OFF_W(l) + 192 * 64


OFF_VE = 640 + 3 * (192 * 64 + 64)
OFF_CE = OFF_VE + 64
OFF_UN = OFF_CE + 1
OFF_CN = OFF_UN + 64
NPARAM = OFF_CN + 1


def rademacher(seed, idx):
    """+-1 from the SplitMix64 finaliser of seed + (idx+1)*golden, as the device draws its probes."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return np.where((z >> np.uint64(63)) != 0, 1.0, -1.0)


class Graph:
    """The pressure operator (diag d, CSR row_ptr/cols/coef of the couplings c_ij) as a graph."""

    def __init__(self, diag, rp, cols, coef):
        n = diag.size
        self.n = n
        self.rp = rp.astype(np.int64)
        self.cols = cols.astype(np.int64)
        deg = np.diff(self.rp)
        self.deg = deg
        self.src = np.repeat(np.arange(n), deg)
        D = diag + np.bincount(self.src, weights=coef, minlength=n)
        self.D = D
        r = diag / D
        self.r = r
        self.rfl = np.maximum(r, R_FLOOR)
        self.a = coef / np.sqrt(D[self.src] * D[self.cols])
        self.coef = coef
        key = {}
        for p, (i, j) in enumerate(zip(self.src, self.cols)):
            key[(i, j)] = p
        self.rev = np.array([key.get((j, i), -1) for i, j in zip(self.src, self.cols)])
        self.x = np.stack([r, np.log10(np.maximum(r, 1e-12)) / 12.0 + 1.0, deg / 8.0, np.ones(n)], axis=1)
        self.f = np.stack([self.a, self.a * deg[self.src] / 6.0, coef / D[self.src], np.ones(self.src.size)], axis=1)


def _t(a):
    return torch.as_tensor(np.asarray(a), dtype=torch.float64)


def forward(theta, g):
    """o_edge, o_node, w (scaled off-diagonal), mdiag, delta; theta a float64 torch vector."""
    th = theta
    Wn = th[OFF_WN:OFF_BN].reshape(4, H)
    bn = th[OFF_BN:OFF_WE]
    We = th[OFF_WE:OFF_BE].reshape(4, H)
    be = th[OFF_BE:OFF_W(0)]
    src = torch.as_tensor(g.src)
    dst = torch.as_tensor(g.cols)
    deg = _t(g.deg)
    h = torch.relu(_t(g.x) @ Wn + bn)
    e = torch.relu(_t(g.f) @ We + be)
    for l in range(3):
        W = th[OFF_W(l):OFF_B(l)].reshape(192, H)
        b = th[OFF_B(l):OFF_B(l) + H]
        X = torch.cat([h[src], h[dst], e], dim=1)
        m = torch.relu(X @ W + b)
        agg = torch.zeros_like(h).index_add_(0, src, m)
        h = h + agg / torch.clamp(deg, min=1.0)[:, None]
        e = m
    oe = e @ th[OFF_VE:OFF_CE] + th[OFF_CE]
    on = h @ th[OFF_UN:OFF_CN] + th[OFF_CN]
    rev = torch.as_tensor(np.where(g.rev >= 0, g.rev, np.arange(g.rev.size)))
    os_ = torch.clamp(0.5 * (oe + oe[rev]), -OCLAMP, OCLAMP)
    w = _t(g.a) * torch.exp(os_)
    delta = _t(g.rfl) * torch.exp(torch.clamp(on, -OCLAMP, OCLAMP))
    mdiag = torch.zeros(g.n, dtype=torch.float64).index_add_(0, src, torch.abs(w)) + delta
    return oe, on, w, mdiag, delta


def probes(g, P, seed, x=None):
    n = g.n
    idx = np.arange(n * P).reshape(P, n)
    v = rademacher(seed, idx)
    if x is not None:
        v[0] = x * np.sqrt(g.D)
    return v


def loss(theta, g, v):
    """Relative spread S / c^2 of Mt At over the probes: S = (1/P) sum_q |Mt At v_q - c v_q|^2 / |v_q|^2,
    c = (1/P) sum_q v_q.Mt At v_q / |v_q|^2, with At = I - N the Jacobi-scaled operator."""
    oe, on, w, mdiag, delta = forward(theta, g)
    src = torch.as_tensor(g.src)
    dst = torch.as_tensor(g.cols)
    V = _t(v)
    a = _t(g.a)
    ys, vs = [], []
    for q in range(V.shape[0]):
        vq = V[q]
        u = vq - torch.zeros(g.n, dtype=torch.float64).index_add_(0, src, a * vq[dst])
        y = mdiag * u + torch.zeros(g.n, dtype=torch.float64).index_add_(0, src, w * u[dst])
        ys.append(y)
        vs.append(vq)
    P = V.shape[0]
    c = sum(torch.dot(vq, y) / torch.dot(vq, vq) for vq, y in zip(vs, ys)) / P
    S = sum(torch.sum((y - c * vq) ** 2) / torch.dot(vq, vq) for vq, y in zip(vs, ys)) / P
    return S / (c * c)


def loss_and_grad(theta_np, g, v):
    th = torch.tensor(np.asarray(theta_np, dtype=np.float64), requires_grad=True)
    L = loss(th, g, v)
    L.backward()
    return float(L.detach()), th.grad.numpy().copy()


def dense_preconditioner(theta_np, g):
    """Mt as a dense matrix (small graphs only)."""
    with torch.no_grad():
        oe, on, w, mdiag, delta = forward(_t(theta_np), g)
    M = np.diag(mdiag.numpy())
    M[g.src, g.cols] += w.numpy()
    return M
