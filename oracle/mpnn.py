"""CPU oracle of the message-passing network (FP64 numpy). Test infrastructure only.

The network (no counterpart in the reference): input layer h0 = relu(X W0 + b0); four layers
a = A h (A: a weighted adjacency, rows summing to one), z = [h | a] W_l + b_l, h <- h + relu(z);
output y = h Wo + bo. With bf16=True every dense product rounds both operands to BF16 (round to
nearest even) and every stored activation to FP32, the points at which the device kernels round;
with bf16=False it is the plain FP64 network, whose gradients central differences check.
"""
import numpy as np

LAYERS = 4
HIDDEN = 64


def bf16(x):
    """Values rounded to FP32 and then to BF16 (round to nearest even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _f32(x, on):
    return np.asarray(x, np.float32).astype(np.float64) if on else x


def _mm(a, w, on):
    return (bf16(a) @ bf16(w)) if on else a @ w


def forward(params, X, A, bf16_on=True):
    """(y, cache). params: W0 (F, 64), b0, W1..W4 (128, 64), b1..b4, Wo (64, m), bo; A: scipy CSR."""
    on = bf16_on
    X = _f32(X, on)
    cache = dict(X=X, h=[], a=[], z=[])
    z0 = _f32(_mm(X, params['W0'], on) + params['b0'], on)
    h = _f32(np.maximum(z0, 0), on)
    cache['z0'] = z0
    for l in range(1, LAYERS + 1):
        a = _f32(A @ h, on)
        z = _f32(_mm(np.hstack([h, a]), params[f'W{l}'], on) + params[f'b{l}'], on)
        cache['h'].append(h)
        cache['a'].append(a)
        cache['z'].append(z)
        h = _f32(h + np.maximum(z, 0), on)
    cache['hout'] = h
    y = _f32(_mm(h, params['Wo'], on) + params['bo'], on)
    return y, cache


def backward(params, cache, A, dy, bf16_on=True):
    """Gradients of sum(dy * y) with respect to every parameter (the device's rounding points)."""
    on = bf16_on
    AT = A.T.tocsr()
    g = {}
    h = cache['hout']
    g['Wo'] = _mm(h.T, dy, on)
    g['bo'] = dy.sum(0)
    dh = _f32(_mm(dy, params['Wo'].T, on), on)
    for l in range(LAYERS, 0, -1):
        hl, al, zl = cache['h'][l - 1], cache['a'][l - 1], cache['z'][l - 1]
        dz = np.where(zl > 0, dh, 0.0)
        g[f'W{l}'] = _mm(np.hstack([hl, al]).T, dz, on)
        g[f'b{l}'] = dz.sum(0)
        D = _f32(_mm(dz, params[f'W{l}'].T, on), on)
        dh = _f32(dh + D[:, :HIDDEN] + AT @ D[:, HIDDEN:], on)
    dz0 = np.where(cache['z0'] > 0, dh, 0.0)
    g['W0'] = _mm(cache['X'].T, dz0, on)
    g['b0'] = dz0.sum(0)
    return g


def init_params(n_in, n_out, seed=0, scale=1.0):
    """Glorot-style random parameters (float64)."""
    rng = np.random.default_rng(seed)
    p = {'W0': rng.normal(0, scale * np.sqrt(2.0 / (n_in + HIDDEN)), (n_in, HIDDEN)),
         'b0': rng.normal(0, 0.1 * scale, HIDDEN)}
    for l in range(1, LAYERS + 1):
        p[f'W{l}'] = rng.normal(0, scale * np.sqrt(2.0 / (3 * HIDDEN)), (2 * HIDDEN, HIDDEN))
        p[f'b{l}'] = rng.normal(0, 0.1 * scale, HIDDEN)
    p['Wo'] = rng.normal(0, scale * np.sqrt(2.0 / (HIDDEN + n_out)), (HIDDEN, n_out))
    p['bo'] = rng.normal(0, 0.1 * scale, n_out)
    return p
