"""CPU oracle of the learned flip scorer (float64 numpy).  Test
infrastructure only: the GPU kernels' forward and backward passes are
checked against it, and it is itself checked by finite differences.

The network has no counterpart in the reference; it follows the layout of
the library's scorer (fg_scorer_layout):

    e_t = relu(x_t W0 + b0)                x_t: factor bits of term t
    h_t = relu(e_t A1 + mean(e) B1 + c1)   mean over the scheme's terms
    g_t = relu(h_t A2 + mean(h) B2 + c2)
    z   = relu([g_a + g_b, g_a * g_b] Wz + bz)
    s   = z . wo[c] + bo[c]
    loss = mean BCE-with-logits(s, y)

With exact=False every GEMM operand is rounded to bfloat16 at the points
where the GPU feeds the MFMA units (weights, activations, gradients), and
everything else stays in high precision, so the two agree up to the FP32
accumulation order.
"""
import numpy as np

D, Z = 128, 64


def bf16(x):
    """Round to bfloat16 (round to nearest even through float32)."""
    a = np.asarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, layout):
    F = int(layout[0])
    off = [int(x) for x in layout[4:16]]
    p = np.asarray(params, dtype=np.float64)
    return {
        "W0": p[off[0]:off[0] + F * D].reshape(F, D), "b0": p[off[1]:off[1] + D],
        "A1": p[off[2]:off[2] + D * D].reshape(D, D), "B1": p[off[3]:off[3] + D * D].reshape(D, D),
        "c1": p[off[4]:off[4] + D],
        "A2": p[off[5]:off[5] + D * D].reshape(D, D), "B2": p[off[6]:off[6] + D * D].reshape(D, D),
        "c2": p[off[7]:off[7] + D],
        "Wz": p[off[8]:off[8] + 2 * D * Z].reshape(2 * D, Z), "bz": p[off[9]:off[9] + Z],
        "wo": p[off[10]:off[10] + 3 * Z].reshape(3, Z), "bo": p[off[11]:off[11] + 3],
    }


def pack(grads, layout, n):
    off = [int(x) for x in layout[4:16]]
    out = np.zeros(n)
    for k, o in zip(["W0", "b0", "A1", "B1", "c1", "A2", "B2", "c2", "Wz", "bz", "wo", "bo"], off):
        g = grads[k].ravel()
        out[o:o + g.size] = g
    return out


def features(words, m, nn):
    """[B, M, F] factor bits of each term (zero beyond the rank)."""
    words = np.asarray(words, dtype=np.uint64)
    B, _, M = words.shape
    cols = []
    for c in range(3):
        w = words[:, c, :]
        cols.append(((w[:, :, None] >> np.arange(nn[c], dtype=np.uint64)) & 1).astype(np.float64))
    X = np.concatenate(cols, axis=2)
    mask = (np.arange(M)[None, :] < np.asarray(m)[:, None]).astype(np.float64)
    return X * mask[:, :, None], mask


def forward(params, layout, words, m, pairs, nn, exact=False):
    r = (lambda a: a) if exact else bf16
    P = unpack(params, layout)
    X, mask = features(words, m, nn)
    n = np.asarray(m, dtype=np.float64)
    E = np.maximum(X @ r(P["W0"]) + P["b0"], 0.0)
    ME = (E * mask[:, :, None]).sum(1) / n[:, None]
    U = r(E) @ r(P["A1"]) + (r(ME) @ r(P["B1"]))[:, None, :] + P["c1"]
    H = np.where((U > 0) & (mask[:, :, None] > 0), U, 0.0)
    MH = (H * mask[:, :, None]).sum(1) / n[:, None]
    V = r(H) @ r(P["A2"]) + (r(MH) @ r(P["B2"]))[:, None, :] + P["c2"]
    G = np.where((V > 0) & (mask[:, :, None] > 0), V, 0.0)
    pairs = np.asarray(pairs)
    bi = np.arange(len(pairs))
    g1, g2 = G[bi, pairs[:, 1]], G[bi, pairs[:, 2]]
    Qp = r(np.concatenate([g1 + g2, g1 * g2], axis=1))
    Zv = np.maximum(Qp @ r(P["Wz"]) + P["bz"], 0.0)
    c = pairs[:, 0]
    s = (Zv * P["wo"][c]).sum(1) + P["bo"][c]
    cache = dict(P=P, X=X, mask=mask, n=n, E=E, ME=ME, H=H, MH=MH, G=G, Qp=Qp, Z=Zv, pairs=pairs, r=r)
    return s, cache


def loss_and_ds(s, y):
    s = np.asarray(s, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    sp = np.where(s > 0, s + np.log1p(np.exp(-np.abs(s))), np.log1p(np.exp(s)))
    loss = float(np.mean(sp - y * s))
    ds = (1.0 / (1.0 + np.exp(-s)) - y) / len(s)
    return loss, ds


def backward(cache, ds):
    r = cache["r"]
    P, X, mask, n = cache["P"], cache["X"], cache["mask"], cache["n"]
    E, ME, H, MH, G, Qp, Zv, pairs = (cache[k] for k in ["E", "ME", "H", "MH", "G", "Qp", "Z", "pairs"])
    c = pairs[:, 0]
    g = {}
    g["wo"] = np.zeros_like(P["wo"])
    g["bo"] = np.zeros_like(P["bo"])
    np.add.at(g["wo"], c, ds[:, None] * Zv)
    np.add.at(g["bo"], c, ds)
    dZ = ds[:, None] * P["wo"][c] * (Zv > 0)
    g["bz"] = dZ.sum(0)
    g["Wz"] = Qp.T @ r(dZ)
    dQ = r(dZ) @ r(P["Wz"]).T
    bi = np.arange(len(pairs))
    dG = np.zeros_like(G)
    g1, g2 = G[bi, pairs[:, 1]], G[bi, pairs[:, 2]]
    np.add.at(dG, (bi, pairs[:, 1]), dQ[:, :D] + dQ[:, D:] * g2)
    np.add.at(dG, (bi, pairs[:, 2]), dQ[:, :D] + dQ[:, D:] * g1)
    dV = dG * (G > 0)
    g["A2"] = np.einsum("btd,bte->de", r(H), r(dV))
    g["c2"] = dV.sum((0, 1))
    SDV = dV.sum(1)
    g["B2"] = r(MH).T @ r(SDV)
    dH = r(dV) @ r(P["A2"]).T
    dMH = r(SDV) @ r(P["B2"]).T
    dU = (dH + (dMH / n[:, None])[:, None, :]) * (H > 0)
    g["A1"] = np.einsum("btd,bte->de", r(E), r(dU))
    g["c1"] = dU.sum((0, 1))
    SDU = dU.sum(1)
    g["B1"] = r(ME).T @ r(SDU)
    dE = r(dU) @ r(P["A1"]).T
    dME = r(SDU) @ r(P["B1"]).T
    dP0 = (dE + (dME / n[:, None])[:, None, :]) * (E > 0) * mask[:, :, None]
    g["W0"] = np.einsum("btf,btd->fd", X, r(dP0))
    g["b0"] = dP0.sum((0, 1))
    return g


def loss_and_grads(params, layout, words, m, pairs, y, nn, exact=False):
    s, cache = forward(params, layout, words, m, pairs, nn, exact)
    loss, ds = loss_and_ds(s, y)
    return loss, pack(backward(cache, ds), layout, len(params)), s


def make_layout(F, M):
    """fg_scorer_layout's 16 values for F input bits and M terms: F D Z M and
    the offsets of W0 b0 A1 B1 c1 A2 B2 c2 Wz bz wo bo (packed in that order)."""
    sizes = [F * D, D, D * D, D * D, D, D * D, D * D, D, 2 * D * Z, Z, 3 * Z, 3]
    off, o = [], 0
    for n in sizes:
        off.append(o)
        o += n
    return np.array([F, D, Z, M] + off, dtype=np.int64), o


def init_params(F, seed, scale=1.0):
    """Random parameters in fg_scorer_layout order (for tests)."""
    layout, n = make_layout(F, 0)
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, n) * scale
    fan = {0: F, 2: D, 3: D, 5: D, 6: D, 8: 2 * D, 10: Z}
    off = list(layout[4:]) + [n]
    for k in range(12):
        a, b = off[k], off[k + 1]
        p[a:b] *= np.sqrt(3.0 / fan[k]) if k in fan else 0.1
    return p
