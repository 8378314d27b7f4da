"""float32 numpy reference of the radiance-cache MLP with the GPU kernels' BF16 roundings:
BF16 weights, BF16 inputs to every product (the encoding and each hidden activation, the
backward gradients), FP32 accumulation and biases.  Layer l: z = W_l h + b_l, ReLU on the
four hidden layers, linear output; loss = mean over samples and channels of the
Huber function (r^2 up to |r| = 10, linear beyond) of r = (y - t) / sqrt(y^2 + 0.01) with the
denominator held constant, or (l2=True) of (y - t)^2."""
import numpy as np

LAYERS = [(32, 64), (64, 64), (64, 64), (64, 64), (64, 3)]
EPS = 0.01
HUBER = 10.0
DTYPE = np.float32  # arithmetic type of forward/backward (float64 when pinning the math)


def bf16(x):
    """round float32 to the nearest bfloat16 (ties to even), returned as float32"""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return ((u + r) & 0xFFFF0000).astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(w):
    out, o = [], 0
    for i, n in LAYERS:
        W = w[o:o + n * i].reshape(n, i)
        o += n * i
        out.append((W, w[o:o + n]))
        o += n
    return out


def encode(pos, d, lo, hi):
    """the 32 encoding dimensions in natural order"""
    pos = np.asarray(pos, np.float32)
    d = np.asarray(d, np.float32)
    ext = (np.asarray(hi, np.float32) - np.asarray(lo, np.float32)).astype(np.float32)
    inv = np.where(ext > 0, np.float32(1.0) / np.where(ext > 0, ext, 1), np.float32(1.0)).astype(np.float32)
    p = ((pos - np.asarray(lo, np.float32)) * inv).astype(np.float32)
    dl = np.sqrt((d * d).sum(1)).astype(np.float32)
    cy = np.clip(d[:, 1] / np.where(dl > 0, dl, 1), -1, 1).astype(np.float32)
    th = (np.arccos(cy) / np.float32(np.pi)).astype(np.float32)
    ph = (np.arctan2(d[:, 2], d[:, 0]) / np.float32(np.pi)).astype(np.float32)
    enc = np.zeros((len(pos), 32), np.float32)
    pi = np.float32(np.pi)
    for e in range(32):
        if e < 24:
            x, r = p[:, e >> 3], e & 7
        else:
            x, r = (th if e < 28 else ph), (e - 24) & 3
        a = (pi * x * np.float32(1 << (r >> 1))).astype(np.float32)
        enc[:, e] = np.cos(a) if r & 1 else np.sin(a)
    return enc


def forward(w, enc):
    """returns (outputs, hs, zs): hs[l] = bf16 input of layer l, zs[l] = pre-activation"""
    hs, zs = [], []
    h = bf16(enc)
    for l, (W, b) in enumerate(unpack(w)):
        hs.append(h)
        z = (h @ bf16(W).T + b).astype(DTYPE)
        zs.append(z)
        h = bf16(np.maximum(z, 0)) if l < 4 else z
    return h, hs, zs


def loss_and_grad(w, enc, target, l2=False):
    y, hs, zs = forward(w, enc)
    n = len(enc)
    if l2:
        e = y - target
        loss = float((e * e).sum() / (3 * n))
        dz = bf16((2 * e / n / 3).astype(DTYPE))
    else:
        sd = np.sqrt(y * y + EPS)
        r = (y - target) / sd
        ar = np.abs(r)
        loss = float(np.where(ar <= HUBER, r * r, HUBER * (2 * ar - HUBER)).sum() / (3 * n))
        dz = bf16((2 * np.clip(r, -HUBER, HUBER) / sd / n / 3).astype(DTYPE))
    grads = []
    layers = unpack(w)
    for l in range(4, -1, -1):
        W, b = layers[l]
        grads.append((dz.T @ hs[l], dz.sum(0)))
        if l == 0:
            break
        dh = (dz @ bf16(W)).astype(DTYPE)
        dz = bf16(np.where(zs[l - 1] > 0, dh, 0).astype(DTYPE))
    grads = grads[::-1]
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads]).astype(DTYPE)
    return loss, flat
