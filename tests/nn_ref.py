"""FP64 numpy restatement of the warm-start network (test infrastructure):
the same graph and parameter layout as seismic_tomography_3d_amd/csrc/
surrogate.hip, used to pin the GPU forward pass and, by central finite
differences, the GPU gradients."""
import numpy as np

# (cin, cout, ksize, resolution divisor, relu)
ARCH = [(6, 16, 3, 1, True), (16, 16, 3, 1, True), (16, 32, 3, 2, True), (32, 32, 3, 2, True),
        (32, 64, 3, 4, True), (64, 64, 3, 8, True), (128, 64, 3, 4, True), (96, 32, 3, 2, True),
        (48, 16, 3, 1, True), (16, 1, 1, 1, False)]


def unpack(p):
    layers, o = [], 0
    for cin, cout, k, _, relu in ARCH:
        t = k ** 3
        W = p[o:o + t * cin * cout].reshape(t * cin, cout)
        o += t * cin * cout
        b = p[o:o + cout]
        o += cout
        layers.append((W, b, k, relu))
    assert o == len(p)
    return layers


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv(X, W, b, k, relu):
    """X: (ns, n, n, n, C) indexed [s, z, y, x, c]; taps t = tx + k (ty + k tz)."""
    ns, n = X.shape[0], X.shape[1]
    cin = X.shape[-1]
    h = k // 2
    Xp = np.pad(X, ((0, 0), (h, h), (h, h), (h, h), (0, 0)))
    out = np.zeros((ns, n, n, n, W.shape[1]))
    for tz in range(k):
        for ty in range(k):
            for tx in range(k):
                t = tx + k * (ty + k * tz)
                sl = Xp[:, tz:tz + n, ty:ty + n, tx:tx + n, :]
                out += sl @ W[t * cin:(t + 1) * cin]
    out += b
    return np.maximum(out, 0.0) if relu else out


def pool(X):
    ns, n = X.shape[0], X.shape[1]
    m = n // 2
    return X.reshape(ns, m, 2, m, 2, m, 2, -1).mean(axis=(2, 4, 6))


def up(X):
    return X.repeat(2, 1).repeat(2, 2).repeat(2, 3)


def forward(p, X, nc):
    L = unpack(np.asarray(p, np.float64))
    x = np.asarray(X, np.float64).reshape(-1, nc, nc, nc, 6)
    a1 = conv(x, *L[0])
    a2 = conv(a1, *L[1])
    a3 = conv(pool(a2), *L[2])
    a4 = conv(a3, *L[3])
    a5 = conv(pool(a4), *L[4])
    a7 = conv(pool(a5), *L[5])
    a8 = conv(np.concatenate([up(a7), a5], -1), *L[6])
    a9 = conv(np.concatenate([up(a8), a4], -1), *L[7])
    a10 = conv(np.concatenate([up(a9), a2], -1), *L[8])
    y = conv(a10, *L[9])
    return y.reshape(x.shape[0], -1)


def loss(p, X, T, M, nc):
    y = forward(p, X, nc)
    T = np.asarray(T, np.float64).reshape(y.shape)
    M = np.asarray(M, np.float64).reshape(y.shape)
    return float((M * (y - T) ** 2).sum() / M.sum())
