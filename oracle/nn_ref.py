"""Oracle (test infrastructure only) of the learned force correction: a numpy restatement of the residual
3-D convolution net and of its parameter gradients, with the BF16 roundings of the device (weights, the
net's input, the stored activations and the gradient tensors; products summed in FP64 here, in FP32 on
the device).  The reference has no learned correction; this restates the library's specification:

    x0 = (log(1 + max(delta, -0.999)), a),  delta = dens * inv_rho - 1
    z1 = W1 * x0 + b1, h1 = relu(z1);  z2 = W2 * h1 + b2, h2 = h1 + relu(z2)
    z3 = W3 * h2 + b3, h3 = h2 + relu(z3);  dphi = s (W4 * h3 + b4)

with periodic 3x3x3 convolutions, W[t][out][in] for taps t = (dz+1) 9 + (dy+1) 3 + (dx+1), and
(W * h)(v) = sum_t W[t] h(v + off_t).  Arrays here are [z, y, x, c].
"""
import numpy as np

PER_LAYER = 27 * 16 * 16 + 16
N_PARAMS = 4 * PER_LAYER
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(x):
    """Round to BF16 (nearest even) and back, through the FP32 bit pattern."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64)
    layers = []
    for l in range(4):
        blk = p[l * PER_LAYER:(l + 1) * PER_LAYER]
        layers.append((blk[:27 * 256].reshape(27, 16, 16).astype(np.float64), blk[27 * 256:].astype(np.float64)))
    return layers


def shift(x, off):
    """x(v + off) for arrays [z, y, x, c] (periodic)."""
    dz, dy, dx = off
    return np.roll(x, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(x, W):
    """sum_t W[t] x(v + off_t): x [z,y,x,ci], W [27, co, ci] -> [z,y,x,co]."""
    out = np.zeros(x.shape[:3] + (W.shape[1],))
    for t, off in enumerate(TAPS):
        out += shift(x, off) @ W[t].T
    return out


def conv_t(g, W):
    """Adjoint of conv: sum_t W[t]^T g(v - off_t)."""
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += shift(g, (-dz, -dy, -dx)) @ W[t]
    return out


def wgrad(g, x):
    """dW[t] = sum_v g(v) x(v + off_t)^T and db = sum_v g(v)."""
    dW = np.zeros((27, g.shape[3], x.shape[3]))
    G = g.reshape(-1, g.shape[3])
    for t, off in enumerate(TAPS):
        dW[t] = G.T @ shift(x, off).reshape(-1, x.shape[3])
    return dW, G.sum(axis=0)


def exact(x):
    """No rounding (the FP64 net, for checks of the chain rule)."""
    return np.asarray(x, np.float64)


def net_input(dens, inv_rho, a, rnd=bf16):
    d = np.asarray(dens, np.float32) * np.float32(inv_rho) - np.float32(1)
    x = np.log1p(np.maximum(d, np.float32(-0.999)).astype(np.float32)).astype(np.float32)
    x0 = np.zeros(d.shape + (16,))
    x0[..., 0] = rnd(x)
    x0[..., 1] = rnd(np.float32(a))
    return x0


def forward(params, dens, inv_rho, a, s_out, rnd=bf16):
    """-> (dphi [z,y,x], cache) for a density mesh dens [z,y,x]; rnd = bf16 (the device's roundings) or
    exact."""
    L = [(rnd(W), b) for W, b in unpack(params)]
    x0 = net_input(dens, inv_rho, a, rnd)
    z1 = conv(x0, L[0][0]) + L[0][1]
    h1 = rnd(np.maximum(z1, 0))
    z2 = conv(h1, L[1][0]) + L[1][1]
    h2 = rnd(h1 + np.maximum(z2, 0))
    z3 = conv(h2, L[2][0]) + L[2][1]
    h3 = rnd(h2 + np.maximum(z3, 0))
    out = conv(h3, L[3][0])[..., 0] + L[3][1][0]
    cache = dict(L=L, x0=x0, z=(rnd(z1), rnd(z2), rnd(z3)), h=(h1, h2, h3), rnd=rnd)
    return s_out * out, cache


def backward(cache, gphi, s_out):
    """Parameter gradient (N_PARAMS) of <gphi, dphi>."""
    L, x0, (z1, z2, z3), (h1, h2, h3) = cache["L"], cache["x0"], cache["z"], cache["h"]
    rnd = cache["rnd"]
    go = np.zeros(gphi.shape + (16,))
    go[..., 0] = rnd(s_out * np.asarray(gphi, np.float32 if rnd is bf16 else np.float64))
    grads = [None] * 4
    grads[3] = wgrad(go, h3)
    g = conv_t(go, L[3][0])
    gh, gz = rnd(g), rnd(np.where(z3 > 0, g, 0))
    grads[2] = wgrad(gz, h2)
    g = gh + conv_t(gz, L[2][0])
    gh, gz = rnd(g), rnd(np.where(z2 > 0, g, 0))
    grads[1] = wgrad(gz, h1)
    g = gh + conv_t(gz, L[1][0])
    gz = rnd(np.where(z1 > 0, g, 0))
    grads[0] = wgrad(gz, x0)
    out = np.zeros(N_PARAMS)
    for l, (dW, db) in enumerate(grads):
        out[l * PER_LAYER:l * PER_LAYER + 27 * 256] = dW.reshape(-1)
        out[l * PER_LAYER + 27 * 256:(l + 1) * PER_LAYER] = db
    return out
