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


# The smoothstep swish a(z) = z S(z), S = t^2 (3 - 2 t), t = clamp(k z + 1/2, 0, 1),
# k the FP16 value next to 1/6.
ACT_K = 0.1666259765625


def act(z):
    t = np.clip(ACT_K * z + 0.5, 0.0, 1.0)
    return z * (t * t * (3.0 - 2.0 * t))


def act_grad(z):
    u = ACT_K * z + 0.5
    t = np.clip(u, 0.0, 1.0)
    dS = np.where((u > 0.0) & (u < 1.0), 6.0 * ACT_K * t * (1.0 - t), 0.0)
    return t * t * (3.0 - 2.0 * t) + z * dS


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
        H = r(act(Z))
        cache["H"].append(H)
        W = r(P["W%d" % layer])
        Z = H @ W.T + P["b%d" % layer]
    return Z, cache


# Inference forms of the network (k_sur_infer): FP16 or BF16 weights and
# activations, FP32 accumulation; the activation in FP16 operations (FP16) or
# in FP32 operations rounded once to BF16 (BF16).


def f16(x):
    """Round to the nearest FP16 (ties to even), returned as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def act_f16(z):
    """a(z) as the FP16 inference kernel evaluates it from the accumulator
    values z: Z = FP16(z), t = clamp(FP16(Z k + 1/2)), u = FP16(t (-2) + 3),
    a = FP16(Z FP16(FP16(t t) u)); each operation rounds once to nearest
    (products and sums of FP16 values are exact in float64 before that)."""
    Z = f16(z)
    t = np.clip(f16(Z * ACT_K + 0.5), 0.0, 1.0)
    u = f16(t * -2.0 + 3.0)
    return f16(Z * f16(f16(t * t) * u))


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def act_bf16(z):
    """a(z) as the BF16 inference kernel evaluates it from the accumulator
    values z: Z = FP32(z), t = clamp(FP32(Z k + 1/2)) (one fused rounding),
    u = FP32(t (-2) + 3), a = FP32(Z FP32(FP32(t t) u)), then rounded to BF16."""
    Z = f32(z)
    t = np.clip(f32(Z * ACT_K + 0.5), 0.0, 1.0)
    u = f32(t * -2.0 + 3.0)
    a = f32(Z * f32(f32(t * t) * u))
    return bf16(a).astype(np.float64)


def forward_inference(p, x, fmt="fp16"):
    """Outputs of the inference kernel's network in the format fmt ("fp16"
    or "bf16"; float64 accumulation where the kernel accumulates in FP32)."""
    r16 = f16 if fmt == "fp16" else (lambda a: bf16(a).astype(np.float64))
    act16 = act_f16 if fmt == "fp16" else act_bf16
    P = unpack(p)
    p32 = {k: np.asarray(v, dtype=np.float32) for k, v in P.items()}
    x = np.asarray(x, dtype=np.float32)
    hi = r16(x).astype(np.float32)
    lo = r16(x - hi)
    X0 = np.concatenate([hi.astype(np.float64), lo], axis=1)
    W0 = r16(p32["W0"])
    Z = X0 @ np.concatenate([W0, W0], axis=1).T + p32["b0"].astype(np.float64)
    for layer in (1, 2, 3, 4):
        H = act16(Z)
        Z = H @ r16(p32["W%d" % layer]).T + p32["b%d" % layer].astype(np.float64)
    return Z


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
        dZ = r(dH * act_grad(c["Z"][layer - 1]))
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
