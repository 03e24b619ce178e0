"""NumPy FP64 restatement of the constitutive surrogate network (forward and
backward), the CPU reference for the GPU training and inference kernels.

With ``rounded=True`` it rounds to BF16 exactly where the GPU kernels do
(weights, inputs split into a BF16 value and a BF16 residual, activations
after SiLU, output errors and back-propagated errors), and accumulates in
FP64 where the GPU accumulates in FP32; with ``rounded=False`` it is the
exact FP64 network, whose gradient the tests pin by finite differences.
"""
import numpy as np

NIN, NOUT, HID = 7, 7, 128
SHAPES = [("W0", (HID, NIN)), ("b0", (HID,)), ("W1", (HID, HID)), ("b1", (HID,)),
          ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)), ("b3", (HID,)),
          ("W4", (NOUT, HID)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name[0] == "W":
            a = np.sqrt(3.0 / shape[1])
            d[name] = rng.uniform(-a, a, size=shape)
        else:
            d[name] = rng.uniform(-0.1, 0.1, size=shape)
    return pack(d)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, rounded=True):
    r = (lambda a: bf16(a).astype(np.float64)) if rounded else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    x = np.asarray(x, dtype=np.float32 if rounded else np.float64)
    if rounded:
        hi = bf16(x)
        lo = bf16(x - hi)
        X0 = np.concatenate([hi, lo], axis=1).astype(np.float64)
        W0 = np.concatenate([r(P["W0"]), r(P["W0"])], axis=1)
    else:
        X0 = x
        W0 = P["W0"]
    cache = {"H": [X0], "Z": []}
    Z = X0 @ W0.T + P["b0"]
    for layer in range(1, 5):
        cache["Z"].append(Z)
        H = r(silu(Z))
        cache["H"].append(H)
        W = r(P["W%d" % layer])
        Z = H @ W.T + P["b%d" % layer]
    return Z, cache


# Inference form of the network (k_sur_infer): FP16 weights and activations,
# layers carrying y = z log2(e) and g = 2 log2(e) SiLU(z),
# g = (y + |y|) + |y| X Q(X), X = 2^-|y|, Q(X) = q0 + ... + q4 X^4 ~ -2 / (1 + X).
SCALE_IN = np.float32(1.4426950408889634)
SCALE_HID = np.float32(0.5)
SCALE_OUT = np.float32(0.34657359027997264)
SILU_Q = (-1.9990234375, 1.970703125, -1.7568359375, 1.126953125, -0.343994140625)


def f16(x):
    """Round to the nearest FP16 (ties to even), returned as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def silu_poly_error(y):
    """|y| X |Q(X) + 2 / (1 + X)| with X = 2^-|y|: the error of g that the
    polynomial adds, exact arithmetic."""
    y = np.abs(np.asarray(y, dtype=np.float64))
    X = 2.0 ** -y
    Q = sum(c * X ** k for k, c in enumerate(SILU_Q))
    return y * X * np.abs(Q + 2.0 / (1.0 + X))


def silu_f16(y):
    """g of the inference kernel for the accumulator values y, in its FP16
    operations: each operation rounds once to nearest (products and sums of
    FP16 values are exact in float64 before that rounding)."""
    y = np.asarray(y, dtype=np.float64)
    X = f16(np.exp2(-np.abs(y)))
    Y = f16(y)
    A = np.abs(Y)
    Q = np.full_like(X, SILU_Q[-1])
    for c in SILU_Q[-2::-1]:
        Q = f16(Q * X + c)
    AX = f16(A * X)
    YA = f16(Y + A)
    return f16(AX * Q + YA)


def forward_inference(p, x):
    """Outputs of the inference kernel's network form (float64 accumulation
    where the kernel accumulates in FP32)."""
    P = unpack(p)
    p32 = {k: np.asarray(v, dtype=np.float32) for k, v in P.items()}
    x = np.asarray(x, dtype=np.float32)
    hi = f16(x).astype(np.float32)
    lo = f16(x - hi)
    X0 = np.concatenate([hi.astype(np.float64), lo], axis=1)
    W0 = f16(p32["W0"] * SCALE_IN)
    Y = X0 @ np.concatenate([W0, W0], axis=1).T + (p32["b0"] * SCALE_IN).astype(np.float64)
    for layer in (1, 2, 3):
        G = silu_f16(Y)
        W = f16(p32["W%d" % layer] * SCALE_HID)
        Y = G @ W.T + (p32["b%d" % layer] * SCALE_IN).astype(np.float64)
    G = silu_f16(Y)
    return G @ f16(p32["W4"] * SCALE_OUT).T + p32["b4"].astype(np.float64)


def loss_and_grad(p, x, t, rounded=True):
    """Mean squared error over all outputs and its gradient (flat vector)."""
    r = (lambda a: bf16(a).astype(np.float64)) if rounded else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    Y, c = forward(p, x, rounded)
    t = np.asarray(t, dtype=np.float64)
    n = Y.shape[0]
    diff = Y - t
    loss = float(np.sum(diff * diff) / (NOUT * n))
    dZ = r(2.0 * diff / (NOUT * n))
    g = {}
    for layer in range(4, -1, -1):
        H = c["H"][layer]
        g["W%d" % layer] = dZ.T @ H
        g["b%d" % layer] = dZ.sum(axis=0)
        if layer == 0:
            break
        W = r(P["W%d" % layer])
        dH = dZ @ W
        dZ = r(dH * silu_grad(c["Z"][layer - 1]))
    if rounded:
        g["W0"] = g["W0"][:, :NIN] + g["W0"][:, NIN:]
    return loss, pack(g)


def neohookean_green(E, lam, mu):
    """PK2 stress (Voigt xx yy zz yz xz xy) and energy of the neo-Hookean law
    from Green-Lagrange strains (n x 6, tensor components)."""
    E = np.asarray(E, dtype=np.float64)
    n = E.shape[0]
    C = np.zeros((n, 3, 3))
    idx = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for k, (i, j) in enumerate(idx):
        C[:, i, j] += 2 * E[:, k]
        if i != j:
            C[:, j, i] += 2 * E[:, k]
    C += np.eye(3)
    Ci = np.linalg.inv(C)
    lnJ = 0.5 * np.log(np.linalg.det(C))
    S = mu * (np.eye(3) - Ci) + (lam * lnJ)[:, None, None] * Ci
    W = 0.5 * lam * lnJ ** 2 + mu * (-lnJ + 0.5 * (np.trace(C, axis1=1, axis2=2) - 3))
    Sv = np.stack([S[:, i, j] for i, j in idx], axis=1)
    return Sv, W


def linear_green(E, lam, mu):
    E = np.asarray(E, dtype=np.float64)
    tr = E[:, 0] + E[:, 1] + E[:, 2]
    S = 2 * mu * E
    S[:, :3] += (lam * tr)[:, None]
    ee = np.sum(E[:, :3] ** 2, axis=1) + 2 * np.sum(E[:, 3:] ** 2, axis=1)
    return S, 0.5 * lam * tr ** 2 + mu * ee
