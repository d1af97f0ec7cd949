"""CPU oracle of the learned Schur-complement correction network.

TEST INFRASTRUCTURE ONLY (imported by ``tests/`` and nothing shipped).

The network has no counterpart in the reference (meeuuw solves Stokes with
SuperLU, meeuuw.py:1832); it is the learned preconditioner stage of the GPU
solver.  This module restates its forward and backward passes in NumPy with
the storage semantics of the GPU kernels, so that the HIP kernels can be
compared with it, and provides an exact float64 mode whose gradient is pinned
by finite differences.

Network (per sample, on the (nelz+1) x (nelx+1) pressure-node grid, NHWC):
    Z1 = conv(X0, W1) + b1,  A1 = relu(Z1)        4 -> 32 channels
    Z2 = conv(A1, W2) + b2,  A2 = relu(Z2)       32 -> 32
    Z3 = conv(A2, W3) + b3,  A3 = relu(Z3)       32 -> 32
    C  = conv(A3, W4) + b4                        32 -> 1
conv is a 3x3 correlation with zero padding: out[y, x] = sum_t W[t] @ X[y+dy, x+dx]
with tap t = 3*(dy+1) + (dx+1); W[t] has shape (c_out, c_in).

GPU storage semantics (``exact=False``): weights are used as bfloat16 copies
of float32 masters, the activations A1..A3 and the back-propagated dZ1..dZ3
are stored as bfloat16 (round to nearest even), X0, C and all accumulations
are float32 or wider.
"""
import numpy as np

HID = 32
CIN = 4
SHAPES = [("W1", (9, HID, CIN)), ("b1", (HID,)), ("W2", (9, HID, HID)), ("b2", (HID,)),
          ("W3", (9, HID, HID)), ("b3", (HID,)), ("W4", (9, 1, HID)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + n]).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _shift(X, dy, dx):
    """S[b, y, x] = X[b, y+dy, x+dx] (zero outside)."""
    B, H, W = X.shape[:3]
    S = np.zeros_like(X)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    S[:, ys:ye, xs:xe] = X[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return S


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(X, W, b=None):
    out = sum(_shift(X, dy, dx) @ W[t].T for t, (dy, dx) in enumerate(TAPS))
    return out if b is None else out + b


def conv_T(D, W):
    """Adjoint of conv in its input: out[p] = sum_t W[t]^T D[p - off_t]."""
    return sum(_shift(D, -dy, -dx) @ W[t] for t, (dy, dx) in enumerate(TAPS))


def weight_grad(D, X):
    """dW[t] = sum_{b,p} D[p] (x) X[p + off_t], shape (9, c_out, c_in)."""
    c_out, c_in = D.shape[-1], X.shape[-1]
    Df = D.reshape(-1, c_out)
    return np.stack([Df.T @ _shift(X, dy, dx).reshape(-1, c_in) for dy, dx in TAPS])


def forward(params, X0, exact=False):
    """X0 (B, H, W, 4) -> (C (B, H, W), cache)."""
    p = unpack(params)
    if exact:
        f = np.float64
        rnd = (lambda a: a)
        Wq = {k: p[k].astype(f) for k in ("W1", "W2", "W3", "W4")}
        X0 = X0.astype(f)
    else:
        f = np.float64
        rnd = bf16
        Wq = {k: bf16(p[k]).astype(f) for k in ("W1", "W2", "W3", "W4")}
        X0 = X0.astype(np.float32).astype(f)
    A1 = rnd(np.maximum(conv(X0, Wq["W1"], p["b1"]), 0)).astype(f)
    A2 = rnd(np.maximum(conv(A1, Wq["W2"], p["b2"]), 0)).astype(f)
    A3 = rnd(np.maximum(conv(A2, Wq["W3"], p["b3"]), 0)).astype(f)
    C = conv(A3, Wq["W4"], p["b4"])[..., 0]
    return C, dict(X0=X0, A1=A1, A2=A2, A3=A3, Wq=Wq, rnd=rnd)


def backward(cache, dC):
    """Gradients of sum(dC * C) with respect to the float32 master parameters
    (the bfloat16 copies pass the gradient straight through)."""
    rnd, Wq = cache["rnd"], cache["Wq"]
    dC = np.asarray(dC, dtype=np.float64)[..., None]
    g = {}
    g["W4"] = weight_grad(dC, cache["A3"])
    g["b4"] = dC.sum(axis=(0, 1, 2))
    dZ3 = rnd(conv_T(dC, Wq["W4"]) * (cache["A3"] > 0)).astype(np.float64)
    g["W3"] = weight_grad(dZ3, cache["A2"])
    g["b3"] = dZ3.sum(axis=(0, 1, 2))
    dZ2 = rnd(conv_T(dZ3, Wq["W3"]) * (cache["A2"] > 0)).astype(np.float64)
    g["W2"] = weight_grad(dZ2, cache["A1"])
    g["b2"] = dZ2.sum(axis=(0, 1, 2))
    dZ1 = rnd(conv_T(dZ2, Wq["W2"]) * (cache["A1"] > 0)).astype(np.float64)
    g["W1"] = weight_grad(dZ1, cache["X0"])
    g["b1"] = dZ1.sum(axis=(0, 1, 2))
    return pack(g)


def init_params(seed):
    """He-uniform weights, zero biases, the output layer scaled by 0.1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = {}
    for name, shp in SHAPES:
        if name[0] == "b":
            d[name] = np.zeros(shp, dtype=np.float32)
        else:
            fan_in = 9 * shp[2]
            lim = np.sqrt(6.0 / fan_in) * (0.1 if name == "W4" else 1.0)
            d[name] = rng.uniform(-lim, lim, size=shp).astype(np.float32)
    return pack(d).astype(np.float32)


def adam(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba, 2015) with bias correction; t counts from 1."""
    m = beta1 * m + (1 - beta1) * grads
    v = beta2 * v + (1 - beta2) * grads * grads
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
