"""CPU restatement (numpy, FP64) of the learned pressure warm start: the 4-level 3-D U-Net, its
two training losses (residual norm, energy) and its gradient.  Test infrastructure only: the tests compare the HIP network
against it, and it is never part of the product path.

The reference has no learned component (its solver is amgcl, src/calc_pressure_3d.cpp:141-166),
so nothing here restates reference code.  The network is the one the library documents in
flux-interpolated-advection_amd/csrc/fia_unet.hpp: channels-last activations, parameters in one
vector, each 3x3x3 convolution W[tap][co][ci] (tap = (dx+1)*9 + (dy+1)*3 + (dz+1)) then bias[co],
the 1x1 output layer w[16], b last.

Pinning: `backward` is checked against central finite differences of `forward` + `loss` in FP64
(directional derivatives, tests/test_oracle_unet.py); the losses are checked against closed forms
(zero residual for the exact solution of the Neumann system; the energy difference equals the
A-norm of the error).  With `bf16=True` the forward and backward
round to BF16 where the GPU stores BF16 (weights, input, activations, pre-activation gradients)
so that the GPU's results can be compared at a tolerance set by FP32 summation order alone.
"""
import numpy as np

# (cin, cout, level, cin_real) of the 14 convolutions, in parameter order
CONVS = [(16, 16, 0, 2), (16, 16, 0, 16), (16, 32, 1, 16), (32, 32, 1, 32), (32, 48, 2, 32), (48, 48, 2, 48),
         (48, 64, 3, 48), (64, 64, 3, 64), (112, 48, 2, 112), (48, 48, 2, 48), (80, 32, 1, 80), (32, 32, 1, 32),
         (48, 16, 0, 48), (16, 16, 0, 16)]


def offsets():
    """(weight offset, bias offset) per convolution, output-layer offset, parameter count."""
    o, out = 0, []
    for cin, cout, _, _ in CONVS:
        out.append((o, o + 27 * cin * cout))
        o += 27 * cin * cout + cout
    return out, o, o + 17


def unpack(params):
    offs, oo, n = offsets()
    assert params.size == n
    Ws, bs = [], []
    for (cin, cout, _, _), (ow, ob) in zip(CONVS, offs):
        Ws.append(params[ow:ow + 27 * cin * cout].reshape(27, cout, cin).astype(np.float64))
        bs.append(params[ob:ob + cout].astype(np.float64))
    return Ws, bs, params[oo:oo + 16].astype(np.float64), float(params[oo + 16])


def bf16(x):
    """Round to BF16 (nearest even on the FP32 bits), returned as FP64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def conv3(x, W, b=None):
    """3x3x3 convolution, stride 1, zero padding: x [X,Y,Z,cin], W [27,cout,cin]."""
    X, Y, Z, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((X, Y, Z, W.shape[1]))
    for t in range(27):
        a, bb, c = t // 9, (t // 3) % 3, t % 3
        out += xp[a:a + X, bb:bb + Y, c:c + Z, :] @ W[t].T
    if b is not None:
        out += b
    return out


def conv3_backward(x, W, dpre):
    """(dW, db, dx) of conv3 for the pre-activation gradient dpre."""
    X, Y, Z, cin = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    D = dpre.reshape(-1, dpre.shape[-1])
    dW = np.zeros_like(W)
    for t in range(27):
        a, bb, c = t // 9, (t // 3) % 3, t % 3
        dW[t] = D.T @ xp[a:a + X, bb:bb + Y, c:c + Z, :].reshape(-1, cin)
        dxp[a:a + X, bb:bb + Y, c:c + Z, :] += dpre @ W[t]
    return dW, D.sum(axis=0), dxp[1:-1, 1:-1, 1:-1, :]


def pool(x):
    X, Y, Z, C = x.shape
    return x.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def up_backward(g):
    X, Y, Z, C = g.shape
    return g.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).sum(axis=(1, 3, 5))


def fluid_flags(solid, shape):
    """Neighbour bits (-x,+x,-y,+y,-z,+z) of fluid cells; 0 on obstacles."""
    fl = np.ones(shape, dtype=bool) if solid is None else (np.asarray(solid) == 0)
    f = np.zeros(shape, dtype=np.int64)
    p = np.pad(fl, 1)
    X, Y, Z = shape
    nb = [p[0:X, 1:Y + 1, 1:Z + 1], p[2:X + 2, 1:Y + 1, 1:Z + 1], p[1:X + 1, 0:Y, 1:Z + 1], p[1:X + 1, 2:Y + 2, 1:Z + 1],
          p[1:X + 1, 1:Y + 1, 0:Z], p[1:X + 1, 1:Y + 1, 2:Z + 2]]
    for k, m in enumerate(nb):
        f |= (m & fl).astype(np.int64) << k
    return f


def laplacian(y, flags):
    """(L y)_c = d_c y_c - sum of y over the fluid neighbours, on cells with neighbours."""
    X, Y, Z = y.shape
    p = np.pad(y, 1)
    nb = [p[0:X, 1:Y + 1, 1:Z + 1], p[2:X + 2, 1:Y + 1, 1:Z + 1], p[1:X + 1, 0:Y, 1:Z + 1], p[1:X + 1, 2:Y + 2, 1:Z + 1],
          p[1:X + 1, 1:Y + 1, 0:Z], p[1:X + 1, 1:Y + 1, 2:Z + 2]]
    out = np.zeros_like(y)
    for k, m in enumerate(nb):
        bit = (flags >> k) & 1
        out += bit * (y - m)
    return np.where((flags & 63) != 0, out, 0.0)


def forward(params, rhs, solid=None, bf=False):
    """Network output y [X,Y,Z] and the cache for backward.  rhs is FP64; solid optional."""
    Ws, bs, wo, bo = unpack(params)
    if bf:
        Ws = [bf16(W) for W in Ws]
    fluid = np.ones(rhs.shape, bool) if solid is None else (np.asarray(solid) == 0)
    n_fl = max(int(fluid.sum()), 1)
    scale = np.sqrt((rhs[fluid] ** 2).sum() / n_fl)
    if not scale > 0:
        scale = 1.0
    b = np.where(fluid, rhs / scale, 0.0)
    x0 = np.zeros(rhs.shape + (16,))
    x0[..., 0] = bf16(b) if bf else b
    x0[..., 1] = (~fluid).astype(np.float64)
    rnd = bf16 if bf else (lambda v: v)
    acts, ins = [None] * 14, [None] * 14

    def conv(i, x):
        ins[i] = x
        acts[i] = rnd(np.maximum(conv3(x, Ws[i], bs[i]), 0.0))
        return acts[i]

    s0 = conv(1, conv(0, x0))
    p0 = rnd(pool(s0))
    s1 = conv(3, conv(2, p0))
    p1 = rnd(pool(s1))
    s2 = conv(5, conv(4, p1))
    p2 = rnd(pool(s2))
    bt = conv(7, conv(6, p2))
    d2 = conv(9, conv(8, np.concatenate([up(bt), s2], axis=-1)))
    d1 = conv(11, conv(10, np.concatenate([up(d2), s1], axis=-1)))
    d0 = conv(13, conv(12, np.concatenate([up(d1), s0], axis=-1)))
    y = d0 @ wo + bo
    if bf:
        y = f32(y)
    return y, {"acts": acts, "ins": ins, "Ws": Ws, "wo": wo, "scale": scale, "b": b, "bf": bf}


def loss(y, cache, flags, mode="residual"):
    """Training loss and dL/dy on the fluid cells, with b = rhs / scale and r = L y - b:
    "residual": |r|^2 / |b|^2;  "energy": (y.L y / 2 - b.y) / |b|^2 = (y.r - b.y) / (2 |b|^2)."""
    fl = (flags & 63) != 0
    b = np.where(fl, cache["b"], 0.0)
    r = np.where(fl, laplacian(y, flags) - b, 0.0)
    if cache["bf"]:
        r = f32(r)
    den = (b ** 2).sum()
    if mode == "residual":
        dy = (2.0 / den) * laplacian(r, flags)
        val = (r ** 2).sum() / den
    else:
        dy = np.where(fl, r / den, 0.0)
        val = 0.5 * ((np.where(fl, y, 0.0) * r).sum() - (b * y).sum()) / den
    if cache["bf"]:
        dy = f32(dy)
    return val, dy


def backward(params, cache, dy):
    """Gradient of the loss with respect to the parameter vector (same layout)."""
    offs, oo, n = offsets()
    g = np.zeros(n)
    acts, ins, Ws = cache["acts"], cache["ins"], cache["Ws"]
    rnd = bf16 if cache["bf"] else (lambda v: v)
    g[oo:oo + 16] = np.tensordot(dy, acts[13], axes=([0, 1, 2], [0, 1, 2]))
    g[oo + 16] = dy.sum()
    gact = [None] * 14
    gact[13] = dy[..., None] * cache["wo"]
    for i in range(13, -1, -1):
        cin, cout, _, _ = CONVS[i]
        dpre = rnd(np.where(acts[i] > 0, gact[i], 0.0))
        dW, db, dx = conv3_backward(ins[i], Ws[i], dpre)
        ow, ob = offs[i]
        g[ow:ow + dW.size] = dW.reshape(-1)
        g[ob:ob + cout] = db
        if i == 0:
            break
        if i in (8, 10, 12):
            ca = {8: 64, 10: 48, 12: 32}[i]
            skip = {8: 5, 10: 3, 12: 1}[i]
            gact[i - 1] = up_backward(dx[..., :ca])
            gact[skip] = dx[..., ca:]
        elif i in (2, 4, 6):
            gp = up(dx) / 8.0
            gact[i - 1] = gact[i - 1] + gp
        else:
            gact[i - 1] = dx
    return g


def init_params(seed=0):
    """He-normal weights (numpy generator; not the library's initial values)."""
    rng = np.random.default_rng(seed)
    offs, oo, n = offsets()
    p = np.zeros(n, dtype=np.float32)
    for (cin, cout, _, creal), (ow, ob) in zip(CONVS, offs):
        W = np.zeros((27, cout, cin))
        W[:, :, :creal] = rng.normal(0.0, np.sqrt(2.0 / (27 * creal)), size=(27, cout, creal))
        p[ow:ow + W.size] = W.reshape(-1)
        p[ob:ob + cout] = rng.normal(0.0, 0.05, size=cout)
    p[oo:oo + 16] = rng.normal(0.0, 0.25, size=16)
    p[oo + 16] = 0.01
    return p
