"""CPU restatement of the learned-ordering MLP (test side): forward and
backward of a 96-H-H-32 ReLU network (H, the hidden width, follows from the
length of the parameter vector: the library's BCC_MLP_HIDDEN) with a mean binary cross-entropy
over 32 bins, evaluated in float64 with BF16 rounding at exactly the points
where the GPU stores BF16 (inputs, weights, hidden activations and the
backpropagated errors fed to the next GEMM)."""
import numpy as np

IN, HID, OUT = 96, 64, 32   # HID: the library's default hidden width


def sizes(hid=HID):
    """Parameter blocks of the flat vector, in the library's order."""
    return [("W1", (IN, hid)), ("b1", (hid,)), ("W2", (hid, hid)), ("b2", (hid,)), ("W3", (hid, OUT)), ("b3", (OUT,))]


def num_params(hid=HID):
    return sum(int(np.prod(s)) for _, s in sizes(hid))


def hidden_of(n):
    """Hidden width of a parameter vector of n entries (H^2 + (IN + OUT + 2) H + OUT = n)."""
    b = IN + OUT + 2
    h = int(round((-b + (b * b + 4 * (n - OUT)) ** 0.5) / 2))
    if num_params(h) != n:
        raise ValueError(f"{n} parameters fit no hidden width")
    return h


def sizes_of(n):
    """Parameter blocks of a vector of n entries."""
    return sizes(hidden_of(n))


SIZES = sizes()


def bf16(x):
    """Round float32 values to BF16 (nearest even) and back to float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def unpack(p):
    out, o = {}, 0
    for name, shape in sizes_of(p.size):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name in ("W1", "b1", "W2", "b2", "W3", "b3")])


def forward_backward(params, x, labels, rounding=True):
    """params: float32 master vector; x: float32 [B, 96] (already BF16 values);
    labels: uint32 [B].  Returns (loss, grads vector float64)."""
    def r(a):
        a = np.asarray(a)
        return bf16(a.astype(np.float32)).astype(np.float64) if rounding else a.astype(np.float64)
    P = unpack(params.astype(np.float64))
    W1, W2, W3 = (r(P[k]) for k in ("W1", "W2", "W3"))
    b1, b2, b3 = P["b1"], P["b2"], P["b3"]
    B = x.shape[0]
    Y = ((labels[:, None] >> np.arange(OUT, dtype=np.uint32)[None, :]) & 1).astype(np.float64)
    X = x.astype(np.float64)
    Z1 = X @ W1 + b1
    H1 = r(np.maximum(Z1, 0))
    Z2 = H1 @ W2 + b2
    H2 = r(np.maximum(Z2, 0))
    Z3 = H2 @ W3 + b3
    loss = np.mean(np.sum(np.maximum(Z3, 0) - Z3 * Y + np.log1p(np.exp(-np.abs(Z3))), axis=1))
    S = 1.0 / (1.0 + np.exp(-Z3))
    G3 = r((S - Y) / B)
    dW3 = H2.T @ G3
    db3 = G3.sum(0)
    D2 = G3 @ W3.T
    G2 = r(np.where(Z2 > 0, D2, 0))
    dW2 = H1.T @ G2
    db2 = G2.sum(0)
    D1 = G2 @ W2.T
    G1 = r(np.where(Z1 > 0, D1, 0))
    dW1 = X.T @ G1
    db1 = G1.sum(0)
    return loss, pack({"W1": dW1, "b1": db1, "W2": dW2, "b2": db2, "W3": dW3, "b3": db3})


def loss_only(params, x, labels):
    P = unpack(params.astype(np.float64))
    B = x.shape[0]
    Y = ((labels[:, None] >> np.arange(OUT, dtype=np.uint32)[None, :]) & 1).astype(np.float64)
    H1 = np.maximum(x @ P["W1"] + P["b1"], 0)
    H2 = np.maximum(H1 @ P["W2"] + P["b2"], 0)
    Z3 = H2 @ P["W3"] + P["b3"]
    return np.mean(np.sum(np.maximum(Z3, 0) - Z3 * Y + np.log1p(np.exp(-np.abs(Z3))), axis=1))


def spread_top4(z, spread=8, taken0=(), far0=None):
    """Order word of one pair's 32 logits: the highest logit first, then each
    time the highest among the bins at least `spread` bins from every bin
    already chosen (any unchosen bin when none is), ties to the lower bin.
    Bins in taken0 are never chosen; bins outside the mask far0 count as
    near a chosen bin from the start."""
    taken = list(taken0)
    far = set(range(32)) if far0 is None else {c for c in range(32) if (far0 >> c) & 1}
    picks = []
    for _ in range(4):
        cand = [c for c in range(32) if c not in taken and c in far]
        if not cand:
            cand = [c for c in range(32) if c not in taken]
        best = cand[0]
        for c in cand[1:]:
            if z[c] > z[best]:
                best = c
        taken.append(best)
        picks.append(best)
        far -= {c for c in range(32) if abs(c - best) < spread}
    return picks[0] | picks[1] << 8 | picks[2] << 16 | picks[3] << 24
