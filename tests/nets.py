"""fp64 numpy restatement of the density/colour MLPs and their backward pass
with the reference's density-gradient injection (src/core/nerf-network.cu:
668-768). Used as the checker of the MFMA kernels; its backward is itself
checked against torch autograd (tests/test_oracle_pins.py,
test_numpy_mlp_backward_matches_autograd).

round16=True restates the fp16 MFMA kernels' arithmetic: the first layer of
each MLP takes its inputs (features, h1, density outputs, SH) as fp16 hi + lo
pairs, exact to about 2^-22 and treated as exact here; h2 and h3 enter their
layers rounded to fp16; in the backward pass every delta and every
activation operand of a weight gradient is rounded to fp16."""
import numpy as np

SH_C = [0.28209479177387814, 0.48860251190291987, 1.0925484305920792, 0.94617469575755997, 0.31539156525251999,
        0.54627421529603959, 0.59004358992664352, 2.8906114426405538, 0.45704579946446572, 0.3731763325901154,
        1.4453057213202769]


def sh4(dir01):
    """dir01: (n, 3) in [0,1] (0.5*d + 0.5) -> (n, 16)."""
    d = 2.0 * np.asarray(dir01, np.float64) - 1.0
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    c = SH_C
    return np.stack([
        np.full_like(x, c[0]), -c[1] * y, c[1] * z, -c[1] * x, c[2] * x * y, -c[2] * y * z, c[3] * zz - c[4],
        -c[2] * x * z, c[5] * (xx - yy), c[6] * y * (-3 * xx + yy), c[7] * x * y * z, c[8] * y * (1 - 5 * zz),
        c[9] * z * (5 * zz - 3), c[8] * x * (1 - 5 * zz), 1.4453057213202769 * z * (xx - yy),
        c[6] * x * (-xx + 3 * yy)], axis=1)


def split(net):
    net = np.asarray(net, np.float64)
    W1 = net[0:2048].reshape(64, 32)
    W2 = net[2048:3072].reshape(16, 64)
    W3 = net[3072:5120].reshape(64, 32)
    W4 = net[5120:9216].reshape(64, 64)
    W5 = net[9216:9408].reshape(3, 64)
    return W1, W2, W3, W4, W5


def _q(a, on):
    """Round to fp16 where the mixed-precision kernel feeds a matrix-core operand."""
    return np.asarray(a, np.float16).astype(np.asarray(a).dtype) if on else a


def forward(net, feat, dir01, round16=False, acc=np.float64, app=None):
    """round16: the fp16 MFMA kernels' operand rounding (module docstring);
    acc: the accumulation type (np.float32 emulates the matrix cores' fp32
    accumulators, np.float64 is exact to working precision).
    app: appearance embedding (w3e [64][16], emb [n_img][16], img [n]): the
    colour MLP's first layer takes [s16 | sh16 | emb[img]] through
    [W3 | w3e], as the reference concatenates them
    (src/core/nerf-network.cu:494-520)."""
    W1, W2, W3, W4, W5 = (w.astype(acc) for w in split(net))
    feat = np.asarray(feat, acc)
    h1 = np.maximum(feat @ W1.T, 0)
    s16 = h1 @ W2.T
    xc = np.concatenate([s16, sh4(dir01).astype(acc)], axis=1)
    z3 = xc @ W3.T
    if app is not None:
        w3e, emb, img = app
        z3 = z3 + np.asarray(emb, acc)[np.asarray(img)] @ np.asarray(w3e, acc).T
    h2 = np.maximum(z3, 0)
    h3 = np.maximum(_q(h2, round16) @ W4.T, 0)
    o = _q(h3, round16) @ W5.T
    rgb = 1.0 / (1.0 + np.exp(-o))
    return dict(h1=h1, s16=s16, xc=xc, h2=h2, h3=h3, o=o, rgb=rgb, feat=feat, round16=round16, acc=acc, app=app)


def backward(net, fw, G_rgb, dL_ds, t, loss_scale, extra=None):
    """Returns (dfeat (n,32), grad_net (9408,)); with an appearance
    embedding in `fw` also (grad_w3e [64][16], grad_emb [n_img][16])."""
    acc = fw.get("acc", np.float64)
    W1, W2, W3, W4, W5 = (w.astype(acc) for w in split(net))
    r = fw.get("round16", False)
    rgb = fw["rgb"]
    dO = _q(np.asarray(G_rgb, acc) * rgb * (1 - rgb), r)
    h1, h2, h3 = _q(fw["h1"], r), _q(fw["h2"], r), _q(fw["h3"], r)
    dW5 = dO.T @ h3
    dH3 = _q((dO @ W5) * (fw["h3"] > 0), r)
    dW4 = dH3.T @ h2
    dz3 = (dH3 @ W4) * (fw["h2"] > 0)
    dH2 = _q(dz3, r)
    dW3 = dH2.T @ _q(fw["xc"], r)
    dXc = dH2 @ W3
    dS = dXc[:, :16].copy()
    g0 = dS[:, 0] + acc(loss_scale) * np.asarray(dL_ds, acc)
    if extra is not None:
        g0 += acc(loss_scale) * np.asarray(extra, acc)
    t = np.asarray(t, acc)
    dS[:, 0] = np.where(t < 1.0, g0 * t * t, g0)
    dS = _q(dS, r)
    dW2 = dS.T @ h1
    dH1 = _q((dS @ W2) * (fw["h1"] > 0), r)
    dW1 = dH1.T @ _q(fw["feat"], r)
    dX = dH1 @ W1
    grad = np.concatenate([dW1.ravel(), dW2.ravel(), dW3.ravel(), dW4.ravel(), dW5.ravel()])
    if fw.get("app") is None:
        return dX, grad
    w3e, emb, img = fw["app"]
    img = np.asarray(img)
    g_w3e = dz3.T @ np.asarray(emb, acc)[img]
    g_emb = np.zeros(np.shape(emb), acc)
    np.add.at(g_emb, img, dz3 @ np.asarray(w3e, acc))
    return dX, grad, g_w3e, g_emb


def xavier(rng, n_out, n_in):
    a = np.sqrt(6.0 / (n_in + n_out))
    return rng.uniform(-a, a, size=(n_out, n_in))


def random_net(seed=0):
    rng = np.random.default_rng(seed)
    return np.concatenate([xavier(rng, 64, 32).ravel(), xavier(rng, 16, 64).ravel(), xavier(rng, 64, 32).ravel(),
                           xavier(rng, 64, 64).ravel(), xavier(rng, 3, 64).ravel()])
