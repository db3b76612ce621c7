"""CPU oracle of the learned smoother -- TEST INFRASTRUCTURE ONLY.

Used by tests/ to check the HIP network kernels (bamrex_amd/csrc/bx_cnn.hip);
the product never imports it.  There is no reference counterpart: the network
is new in this build (the reference solves the pressure equation with plain
AMReX MLMG, Source/IMEX/IMEX_O1.cpp:616-640).  What pins this oracle:
  * forward: exact-arithmetic cases (small-integer inputs and weights make
    every BF16 rounding and FP32 sum exact), homogeneity N(c x) = c N(x) for
    powers of two, identity network;
  * backward: central finite differences of the same network evaluated in
    float64 without BF16 rounding (tests/test_helmholtz_oracle.py).

Network: bias-free 3x3 convolutions 1->16->16->16->1, ReLU after the first
three.  Parameters flat: W1[9][16], W2[9][16][16], W3[9][16][16], W4[9][16]
([tap][out][in], tap = (dy+1)*3 + dx+1 for the neighbour (x+dx, y+dy)).
"""
import numpy as np

C = 16
NW1, NW2, NW3, NW4 = 9 * C, 9 * C * C, 9 * C * C, 9 * C
NPARAMS = NW1 + NW2 + NW3 + NW4


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(p):
    p = np.asarray(p)
    o = 0
    W1 = p[o:o + NW1].reshape(9, C); o += NW1
    W2 = p[o:o + NW2].reshape(9, C, C); o += NW2
    W3 = p[o:o + NW3].reshape(9, C, C); o += NW3
    W4 = p[o:o + NW4].reshape(9, C)
    return W1, W2, W3, W4


def _shift_img(a, dy, dx, periodic, chan_last):
    """s[..., y, x] = a[..., y+dy, x+dx] (zero outside unless periodic); the
    spatial axes are the last two, or -3 and -2 when channels are last."""
    ya, xa = (-3, -2) if chan_last else (-2, -1)
    if periodic:
        return np.roll(np.roll(a, -dy, axis=ya), -dx, axis=xa)
    out = np.zeros_like(a)
    H, W = a.shape[ya], a.shape[xa]
    ys = slice(max(0, -dy), min(H, H - dy))
    xs = slice(max(0, -dx), min(W, W - dx))
    yd = slice(max(0, dy), min(H, H + dy))
    xd = slice(max(0, dx), min(W, W + dx))
    if chan_last:
        out[..., ys, xs, :] = a[..., yd, xd, :]
    else:
        out[..., ys, xs] = a[..., yd, xd]
    return out


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def forward(params, x, periodic=False, quantize=True):
    """y = N(x) for x (nimg, H, W).  quantize=True emulates the device:
    BF16 weights/activations, FP32 sums in the device's order for layers 1
    and 4 (bit-exact) and a float64 sum rounded to FP32 for the MFMA layers.
    quantize=False: plain float64 network.  Returns (y, cache)."""
    W1, W2, W3, W4 = unpack(params)
    if quantize:
        q = lambda v: bf16(np.asarray(v, np.float32))
        W1, W2, W3, W4 = (bf16(w.astype(np.float32)) for w in (W1, W2, W3, W4))
        x = q(x)
        # layer 1: s += W1[t][n] * v[t] in float32, t = 0..8
        s = np.zeros(x.shape + (C,), np.float32)
        for t, (dy, dx) in enumerate(TAPS):
            v = _shift_img(x, dy, dx, periodic, False)
            s = (s + (v[..., None] * W1[t][None, None, None, :]).astype(np.float32)).astype(np.float32)
        h1 = q(np.maximum(s, 0))
        hs = [h1]
        for W in (W2, W3):
            acc = np.zeros(x.shape + (C,), np.float64)
            for t, (dy, dx) in enumerate(TAPS):
                v = _shift_img(hs[-1], dy, dx, periodic, True).astype(np.float64)
                acc += v @ W[t].astype(np.float64).T
            hs.append(q(np.maximum(acc.astype(np.float32), 0)))
        h3 = hs[-1]
        # layer 4: s += W4[t][k] * h3[q][k] in float32, t outer, k inner
        s = np.zeros(x.shape, np.float32)
        for t, (dy, dx) in enumerate(TAPS):
            v = _shift_img(h3, dy, dx, periodic, True)
            for k in range(C):
                s = (s + (W4[t][k] * v[..., k]).astype(np.float32)).astype(np.float32)
        y = s
        cache = dict(x=x, h1=h1, h2=hs[1], h3=h3, W=(W1, W2, W3, W4))
        return y, cache
    x = np.asarray(x, np.float64)
    a1 = np.zeros(x.shape + (C,))
    for t, (dy, dx) in enumerate(TAPS):
        a1 += _shift_img(x, dy, dx, periodic, False)[..., None] * W1[t]
    h1 = np.maximum(a1, 0)
    hs = [h1]
    for W in (W2, W3):
        acc = np.zeros(x.shape + (C,))
        for t, (dy, dx) in enumerate(TAPS):
            acc += _shift_img(hs[-1], dy, dx, periodic, True) @ W[t].T
        hs.append(np.maximum(acc, 0))
    y = np.zeros(x.shape)
    for t, (dy, dx) in enumerate(TAPS):
        y += _shift_img(hs[-1], dy, dx, periodic, True) @ W4[t]
    return y, dict(x=x, h1=h1, h2=hs[1], h3=hs[2], W=(W1, W2, W3, W4))


def backward(cache, gy, periodic=False):
    """dL/dparams (float64, flat) given dL/dy, treating the BF16 roundings of
    the forward as identity (straight-through), as the device does."""
    W1, W2, W3, W4 = (np.asarray(w, np.float64) for w in cache["W"])
    x, h1, h2, h3 = (np.asarray(cache[k], np.float64) for k in ("x", "h1", "h2", "h3"))
    gy = np.asarray(gy, np.float64)
    dW4 = np.zeros((9, C))
    g3 = np.zeros(h3.shape)
    for t, (dy, dx) in enumerate(TAPS):
        dW4[t] = gy.reshape(-1) @ _shift_img(h3, dy, dx, periodic, True).reshape(-1, C)
        g3 += _shift_img(gy, -dy, -dx, periodic, False)[..., None] * W4[t]
    g3 *= h3 > 0

    def layer(gout, hin, W):
        dW = np.zeros((9, C, C))
        gin = np.zeros(hin.shape)
        for t, (dy, dx) in enumerate(TAPS):
            dW[t] = gout.reshape(-1, C).T @ _shift_img(hin, dy, dx, periodic, True).reshape(-1, C)
            gin += _shift_img(gout, -dy, -dx, periodic, True) @ W[t]
        return dW, gin * (hin > 0)

    dW3, g2 = layer(g3, h2, W3)
    dW2, g1 = layer(g2, h1, W2)
    dW1 = np.zeros((9, C))
    for t, (dy, dx) in enumerate(TAPS):
        dW1[t] = _shift_img(x, dy, dx, periodic, False).reshape(-1) @ g1.reshape(-1, C)
    return np.concatenate([dW1.ravel(), dW2.ravel(), dW3.ravel(), dW4.ravel()])


# ------------------------------------------------------- training loss
def op_diag(a, bX, bY, alpha, beta, dx, dy):
    ny, nx = a.shape
    dhx, dhy = beta / (dx * dx), beta / (dy * dy)
    return alpha * a + dhx * (bX[:, :nx] + bX[:, 1:]) + dhy * (bY[:ny] + bY[1:])


def _gather(F, f, x0, y0, cs):
    ny, nx = F.shape[-2:]
    ys = (y0 + np.arange(cs)) % ny
    xs = (x0 + np.arange(cs)) % nx
    return F[f][np.ix_(ys, xs)] if F.ndim == 3 else F[np.ix_(ys, xs)]


def _apply_local(u, a, bxl, bxr, byl, byr, alpha, dhx, dhy):
    """A u on a crop with u = 0 outside it (crop-local coefficient arrays)."""
    def at(dy, dx):
        return _shift_img(u, dy, dx, False, False)
    uc = u
    return (alpha * a * uc - dhx * (bxr * (at(0, 1) - uc) - bxl * (uc - at(0, -1)))
            - dhy * (byr * (at(1, 0) - uc) - byl * (uc - at(-1, 0))))


def crop_loss_grad(params, R, crops, cs, a, bX, bY, alpha, beta, dx, dy, border=5):
    """Loss and gradient of the device's training objective (float64 except
    the emulated BF16 forward):
      L = mean_b ||m (r_b - A e_b)||^2 / n_b,
      n_b = max(||m r_b||^2, |m| sigma_b^2)  (floored: a crop whose residual
            sits in its border, which the network sees, has a near-zero
            masked norm),
      e_b = sigma_b * N(r_b / sigma_b) / D,  sigma_b = rms of the crop."""
    ny, nx = a.shape
    D = op_diag(a, bX, bY, alpha, beta, dx, dy)
    dhx, dhy = beta / (dx * dx), beta / (dy * dy)
    bxl_f, bxr_f = bX[:, :nx], bX[:, 1:]
    byl_f, byr_f = bY[:ny], bY[1:]
    B = len(crops)
    xs = np.empty((B, cs, cs), np.float32)
    meta = []
    for b, (f, x0, y0) in enumerate(crops):
        r = _gather(R, f, x0, y0, cs)
        sig = np.sqrt(np.mean(r * r))
        xs[b] = (r * (1.0 / sig if sig > 0 else 0.0)).astype(np.float32)
        meta.append((r, sig))
    y, cache = forward(params, xs, periodic=False, quantize=True)
    loss = 0.0
    gy = np.zeros((B, cs, cs))
    mask = np.zeros((cs, cs), bool)
    mask[border:cs - border, border:cs - border] = True
    for b, (f, x0, y0) in enumerate(crops):
        r, sig = meta[b]
        Dl = _gather(D[None], 0, x0, y0, cs)
        al = _gather(a[None], 0, x0, y0, cs)
        c = [_gather(z[None], 0, x0, y0, cs) for z in (bxl_f, bxr_f, byl_f, byr_f)]
        e = sig * y[b].astype(np.float64) / Dl
        res = np.where(mask, r - _apply_local(e, al, *c, alpha, dhx, dhy), 0.0)
        rn2 = max(np.sum(np.where(mask, r * r, 0.0)), np.sum(r * r) * (mask.sum() / float(cs * cs)))
        rn2 = rn2 if rn2 > 0 else 1.0
        loss += np.sum(res * res) / rn2
        ar = _apply_local(res, al, *c, alpha, dhx, dhy)
        gy[b] = (-2.0 * ar / rn2) * sig / Dl / B
    g = backward(cache, gy, periodic=False)
    return loss / B, g
