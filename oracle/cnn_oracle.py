"""FP64 numpy restatement of the learned-proposal CNN -- TEST INFRASTRUCTURE.

Same architecture, parameter layout, output head and loss as the HIP kernels
(thx_cnn.hip): conv 1->16 5x5/2, 16->32 3x3/2, 32->64 3x3/2, 64->128 3x3/2
(each + ReLU), fc 2048->384 + ReLU, fc 384->6C+3.  The backward pass is
derived by hand (im2col form); tests pin it against finite differences and
torch autograd.  There is no reference counterpart: the reference has no
learned proposal.
"""
import numpy as np

CONV = [(1, 64, 16, 32, 5, 2, 2), (16, 32, 32, 16, 3, 2, 1), (32, 16, 64, 8, 3, 2, 1),
        (64, 8, 128, 4, 3, 2, 1)]
FC_IN, FC_HID = 2048, 384


def layout(C):
    nout = 6 * C + 3
    shapes = []
    for ci, hi, co, ho, ks, st, pd in CONV:
        shapes += [(co, ci, ks, ks), (co,)]
    shapes += [(FC_HID, FC_IN), (FC_HID,), (nout, FC_HID), (nout,)]
    return shapes


def unpack(p, C):
    out, o = [], 0
    for s in layout(C):
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s).astype(np.float64))
        o += n
    assert o == p.size
    return out


def pack(arrs):
    return np.concatenate([a.ravel() for a in arrs])


def _im2col(x, ks, st, pd, ho):
    B, ci, hi, _ = x.shape
    xp = np.zeros((B, ci, hi + 2 * pd + st, hi + 2 * pd + st))
    xp[:, :, pd:pd + hi, pd:pd + hi] = x
    cols = np.zeros((B, ci, ks, ks, ho, ho))
    for ky in range(ks):
        for kx in range(ks):
            cols[:, :, ky, kx] = xp[:, :, ky:ky + st * ho:st, kx:kx + st * ho:st]
    return cols.reshape(B, ci * ks * ks, ho * ho)


def _col2im(cols, shape, ks, st, pd, ho):
    B, ci, hi, _ = shape
    xp = np.zeros((B, ci, hi + 2 * pd + st, hi + 2 * pd + st))
    c = cols.reshape(B, ci, ks, ks, ho, ho)
    for ky in range(ks):
        for kx in range(ks):
            xp[:, :, ky:ky + st * ho:st, kx:kx + st * ho:st] += c[:, :, ky, kx]
    return xp[:, :, pd:pd + hi, pd:pd + hi]


def forward(p, x, C):
    W = unpack(p, C)
    acts = [x.astype(np.float64)]
    cache = []
    h = acts[0]
    for l, (ci, hi, co, ho, ks, st, pd) in enumerate(CONV):
        cols = _im2col(h, ks, st, pd, ho)
        w = W[2 * l].reshape(co, -1)
        y = np.einsum("ok,bkn->bon", w, cols) + W[2 * l + 1][None, :, None]
        h = np.maximum(y, 0).reshape(-1, co, ho, ho)
        cache.append(cols)
        acts.append(h)
    f = h.reshape(h.shape[0], -1)
    z = np.maximum(f @ W[8].T + W[9], 0)
    out = z @ W[10].T + W[11]
    return out, (acts, cache, f, z)


def _sigm(a):
    return 1.0 / (1.0 + np.exp(-a))


def head(out, C):
    """Mixture parameters (C x (m, k, pi), mu, sigma) from raw outputs."""
    raw = out[:, :6 * C].reshape(-1, C, 6)
    m = raw[:, :, :4] / (np.linalg.norm(raw[:, :, :4], axis=2, keepdims=True) + 1e-12)
    k = np.exp(-(1.0 + 6.0 * _sigm(raw[:, :, 4])))
    lg = raw[:, :, 5] - raw[:, :, 5].max(1, keepdims=True)
    pi = np.exp(lg) / np.exp(lg).sum(1, keepdims=True)
    mu = out[:, 6 * C:6 * C + 2]
    sig = np.clip(np.exp(out[:, 6 * C + 2]), 0.2, 20.0)
    return m, k, pi, mu, sig


def loss_and_dout(out, q, t, C):
    """Mean loss and its gradient w.r.t. the raw outputs."""
    B = out.shape[0]
    raw = out[:, :6 * C].reshape(B, C, 6)
    r4 = raw[:, :, :4]
    n = np.linalg.norm(r4, axis=2, keepdims=True) + 1e-12
    m = r4 / n
    d = np.einsum("bca,ba->bc", m, q)
    sg = _sigm(raw[:, :, 4])
    k = np.exp(-(1.0 + 6.0 * sg))
    s = d * d + (1 - d * d) / k
    logA = -np.log(2 * np.pi ** 2) - 1.5 * np.log(k) - 2 * np.log(s)
    lgt = raw[:, :, 5]
    logpi = lgt - lgt.max(1, keepdims=True)
    logpi = logpi - np.log(np.exp(logpi).sum(1, keepdims=True))
    lc = logpi + logA
    lmax = lc.max(1, keepdims=True)
    lse = lmax[:, 0] + np.log(np.exp(lc - lmax).sum(1))
    gam = np.exp(lc - lse[:, None])
    dlogA_dd = -4 * d * (1 - 1 / k) / s
    dlogA_dk = -1.5 / k + 2 * (1 - d * d) / (k * k * s)
    g = np.zeros_like(out)
    gr = np.zeros((B, C, 6))
    gr[:, :, :4] = (-gam * dlogA_dd)[:, :, None] * (q[:, None, :] - d[:, :, None] * m) / n
    gr[:, :, 4] = -gam * dlogA_dk * (-k * 6 * sg * (1 - sg))
    gr[:, :, 5] = np.exp(logpi) - gam
    g[:, :6 * C] = gr.reshape(B, -1)
    bs = out[:, 6 * C + 2]
    e = np.exp(bs)
    sig = np.clip(e, 0.2, 20.0)
    dx = t[:, 0] - out[:, 6 * C]
    dy = t[:, 1] - out[:, 6 * C + 1]
    r2 = dx * dx + dy * dy
    lt = np.log(2 * np.pi) + 2 * np.log(sig) + r2 / (2 * sig * sig)
    g[:, 6 * C] = -dx / sig ** 2
    g[:, 6 * C + 1] = -dy / sig ** 2
    g[:, 6 * C + 2] = np.where((e < 0.2) | (e > 20.0), 0.0, 2 - r2 / sig ** 2)
    loss = -lse + lt
    return loss.mean(), g / B


def backward(p, cache_all, dout, C):
    W = unpack(p, C)
    acts, cache, f, z = cache_all
    grads = [None] * len(W)
    grads[10] = dout.T @ z
    grads[11] = dout.sum(0)
    dz = (dout @ W[10]) * (z > 0)
    grads[8] = dz.T @ f
    grads[9] = dz.sum(0)
    dh = (dz @ W[8]).reshape(acts[4].shape)
    for l in range(3, -1, -1):
        ci, hi, co, ho, ks, st, pd = CONV[l]
        gy = (dh * (acts[l + 1] > 0)).reshape(-1, co, ho * ho)
        cols = cache[l]
        grads[2 * l] = np.einsum("bon,bkn->ok", gy, cols).reshape(co, ci, ks, ks)
        grads[2 * l + 1] = gy.sum((0, 2))
        if l > 0:
            dcols = np.einsum("ok,bon->bkn", W[2 * l].reshape(co, -1), gy)
            dh = _col2im(dcols, acts[l].shape, ks, st, pd, ho)
    return pack(grads)


def loss_grad(p, x, q, t, C):
    out, cache = forward(p, x, C)
    loss, dout = loss_and_dout(out, q, t, C)
    return loss, backward(p, cache, dout, C), out
