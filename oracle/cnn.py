"""FP64 restatement of the learned warm-start network (test infrastructure
only: the product runs imp_cnn_* in libimpatient.so).

The network (impatient-mri_amd/csrc/cnn.hip) has six 3x3 convolutions on
H x W x 32 NHWC images with zero padding: layers 0-4 conv + bias + ReLU,
layer 5 conv + bias; x = z + s * net(z / s) with s = rms(z).  Parameters are
flat, per layer [9 taps][32 co][32 ci] weights (tap = (dy+1)*3 + (dx+1))
then 32 biases.  With bf16=True the functions round to bfloat16 (round to
nearest even) exactly where the GPU stores bf16: the weights, the input
features, every post-ReLU activation and every backward gradient image; the
arithmetic itself is FP64.  With bf16=False everything is FP64, which is
what the finite-difference pin of the backward pass uses.

Not in the reference (its CG starts from zero, recon.cu:1871).
"""
import numpy as np

CH = 32
NLAYER = 6
WSZ = 9 * CH * CH
PSZ = WSZ + CH
NPARAM = NLAYER * PSZ
CIO = 2


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(x):
    return np.asarray(x, np.float64)


def unpack(params):
    p = np.asarray(params, np.float64).reshape(NLAYER, PSZ)
    W = p[:, :WSZ].reshape(NLAYER, 9, CH, CH)
    b = p[:, WSZ:]
    return W, b


def conv3x3(a, W, b=None):
    """out[y, x, co] = b[co] + sum_{tap, ci} a[y+dy, x+dx, ci] W[tap, co, ci]."""
    H, Wd, _ = a.shape
    ap = np.pad(a, ((1, 1), (1, 1), (0, 0)))
    out = np.zeros((H, Wd, W.shape[1]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += ap[1 + dy:1 + dy + H, 1 + dx:1 + dx + Wd, :] @ W[tap].T
    if b is not None:
        out += b
    return out


def conv3x3_bwd_data(g, W):
    """dL/da[p, ci] = sum_{tap, co} g[p - off(tap), co] W[tap, co, ci]."""
    H, Wd, _ = g.shape
    gp = np.pad(g, ((1, 1), (1, 1), (0, 0)))
    out = np.zeros((H, Wd, W.shape[2]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += gp[1 - dy:1 - dy + H, 1 - dx:1 - dx + Wd, :] @ W[tap]
    return out


def wgrad(a, g):
    """dW[tap, co, ci] = sum_p g[p, co] a[p + off(tap), ci]; db[co] = sum_p g[p, co]."""
    H, Wd, _ = a.shape
    ap = np.pad(a, ((1, 1), (1, 1), (0, 0)))
    dW = np.zeros((9, CH, CH))
    G = g.reshape(-1, CH)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        A = ap[1 + dy:1 + dy + H, 1 + dx:1 + dx + Wd, :].reshape(-1, CH)
        dW[tap] = G.T @ A
    return dW, G.sum(axis=0)


def _features(z, rnd):
    z = np.asarray(z, np.complex128)
    s = np.sqrt(np.mean(np.abs(z) ** 2))
    u = z / s
    a0 = np.zeros(z.shape + (CH,))
    a0[..., 0] = rnd(u.real)
    a0[..., 1] = rnd(u.imag)
    return s, u, a0


def forward(params, z, bf16_round=True):
    """x = z + s * net(z / s) for z [H, W] complex; returns (x, cache)."""
    rnd = bf16 if bf16_round else _id
    W, b = unpack(params)
    Wr = rnd(W)
    s, u, a = _features(z, rnd)
    acts = [a]
    for layer in range(NLAYER - 1):
        a = rnd(np.maximum(conv3x3(a, Wr[layer], b[layer]), 0.0))
        acts.append(a)
    net = conv3x3(a, Wr[NLAYER - 1], b[NLAYER - 1])
    x = np.asarray(z, np.complex128) + s * (net[..., 0] + 1j * net[..., 1])
    return x, dict(s=s, u=u, acts=acts, net=net, Wr=Wr)


def loss_and_grad(params, z, t, bf16_round=True):
    """loss = 0.5 |x - t|^2 / |t|^2 and its gradient (flat, NPARAM)."""
    rnd = bf16 if bf16_round else _id
    x, c = forward(params, z, bf16_round)
    s, acts, Wr = c["s"], c["acts"], c["Wr"]
    t = np.asarray(t, np.complex128)
    e = (x - t) / s                      # = u + net - t / s
    T = np.sum(np.abs(t / s) ** 2)
    loss = 0.5 * np.sum(np.abs(e) ** 2) / T
    g = np.zeros(z.shape + (CH,))
    g[..., 0] = rnd(e.real / T)
    g[..., 1] = rnd(e.imag / T)
    grad = np.zeros((NLAYER, PSZ))
    for layer in range(NLAYER - 1, -1, -1):
        dW, db = wgrad(acts[layer], g)
        grad[layer, :WSZ] = dW.reshape(-1)
        grad[layer, WSZ:] = db
        if layer == 0:
            break
        ga = conv3x3_bwd_data(g, Wr[layer])
        g = rnd(ga * (acts[layer] > 0))
    # entries the GPU keeps at zero (layer 0 reads channels < CIO, layer 5
    # writes channels < CIO)
    grad[0, :WSZ].reshape(9, CH, CH)[:, :, CIO:] = 0.0
    grad[NLAYER - 1, :WSZ].reshape(9, CH, CH)[:, CIO:, :] = 0.0
    grad[NLAYER - 1, WSZ + CIO:] = 0.0
    return loss, grad.reshape(-1)


def trainable_mask():
    m = np.ones((NLAYER, PSZ), bool)
    m[0, :WSZ].reshape(9, CH, CH)[:, :, CIO:] = False
    m[NLAYER - 1, :WSZ].reshape(9, CH, CH)[:, CIO:, :] = False
    m[NLAYER - 1, WSZ + CIO:] = False
    return m.reshape(-1)


def random_params(seed, scale=1.0):
    """He-scaled random parameters on the trainable entries, nonzero biases
    and output layer (so every gradient path is exercised)."""
    rng = np.random.default_rng(seed)
    p = np.zeros((NLAYER, PSZ))
    for layer in range(NLAYER):
        fan = 9 * (CIO if layer == 0 else CH)
        p[layer, :WSZ] = rng.standard_normal(WSZ) * np.sqrt(2.0 / fan) * scale
        p[layer, WSZ:] = rng.standard_normal(CH) * 0.05
    p = p.reshape(-1)
    p[~trainable_mask()] = 0.0
    return p
