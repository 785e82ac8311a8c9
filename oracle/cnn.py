"""CPU oracle of the learned potential-product network (TEST INFRASTRUCTURE ONLY).

Plain numpy restatement of the network the HIP kernels implement, used by the
tests to check the GPU forward pass and the GPU gradients. The network has no
reference counterpart (the reference has no learned component); its maths:

  u[c]    CI real input fields on the grid (CI, n2, n1, n0), x fastest; zero outside
  h[a]    = sum_t sum_c W1[t, c, a] u[c](v + t)        t in {-1,0,1}^3, a < C
  m[a]    = P1[a](v) h[a]                              P1: fixed positional fields
  o[b]    = sum_t sum_a W2[t, a, b] m[a](v + t)        b < C
  out(v)  = sum_b P2[b](v) o[b](v)
  loss    = sum_v (out - target)^2 / sum_v target^2

Tap index t = (dz+1)*9 + (dy+1)*3 + (dx+1). Gradients by the chain rule; they
are pinned by central finite differences in tests/test_learned.py.

The input fields of the preconditioner's network are y and its first moments
G_d = X_d y - x_d * (M y) (features()).
"""
import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(u, dz, dy, dx):
    """s(v) = u(v + (dz, dy, dx)), zero outside the grid (u: (..., n2, n1, n0))."""
    out = np.zeros_like(u)
    n2, n1, n0 = u.shape[-3:]
    sz = slice(max(0, -dz), n2 - max(0, dz))
    sy = slice(max(0, -dy), n1 - max(0, dy))
    sx = slice(max(0, -dx), n0 - max(0, dx))
    tz = slice(max(0, dz), n2 - max(0, -dz))
    ty = slice(max(0, dy), n1 - max(0, -dy))
    tx = slice(max(0, dx), n0 - max(0, -dx))
    out[..., sz, sy, sx] = u[..., tz, ty, tx]
    return out


def features(oracle, y, greville):
    """[y, G_x, G_y, G_z] for a complex coefficient vector y (N,), G_d = X_d y - x_d M y
    with the oracle's mass and position operators; greville: coordinates per axis."""
    n0, n1, n2 = oracle.n
    My = oracle.term(0, y)
    u = [y]
    idx = np.arange(oracle.N)
    ijk = (idx % n0, (idx // n0) % n1, idx // (n0 * n1))
    for d in range(3):
        if d < oracle.desc["dim"]:
            x = np.asarray(greville[d])[ijk[d]]
            u.append(oracle.term(4 + d, y) - x * My)
        else:
            u.append(np.zeros_like(y))
    return np.array(u)


def forward(u, W1, W2, P1, P2):
    """u: (CI, n2, n1, n0); W1: (27, CI, C); W2: (27, C, C); P1, P2: (C, n2, n1, n0)."""
    C = W1.shape[2]
    h = np.zeros((C,) + u.shape[1:])
    for t, (dz, dy, dx) in enumerate(TAPS):
        s = shift(u, dz, dy, dx)
        h += np.tensordot(W1[t], s, axes=([0], [0]))
    m = P1 * h
    o = np.zeros_like(m)
    for t, (dz, dy, dx) in enumerate(TAPS):
        s = shift(m, dz, dy, dx)
        o += np.tensordot(W2[t], s, axes=([0], [0]))
    out = np.sum(P2 * o, axis=0)
    return out, (h, m, o)


def loss_and_grad(u, target, W1, W2, P1, P2):
    C = W1.shape[2]
    CI = W1.shape[1]
    out, (h, m, o) = forward(u, W1, W2, P1, P2)
    T2 = np.sum(target ** 2)
    diff = out - target
    loss = np.sum(diff ** 2) / T2
    dout = 2.0 * diff / T2
    do = P2 * dout[None]
    dW2 = np.zeros_like(W2)
    dm = np.zeros_like(m)
    for t, (dz, dy, dx) in enumerate(TAPS):
        s = shift(m, dz, dy, dx)
        dW2[t] = s.reshape(C, -1) @ do.reshape(C, -1).T
        # m(v + t) feeds o(v): dm(w) += sum_b W2[t, a, b] do(w - t)
        dm += shift(np.tensordot(W2[t], do, axes=([1], [0])), -dz, -dy, -dx)
    dh = P1 * dm
    dW1 = np.zeros_like(W1)
    for t, (dz, dy, dx) in enumerate(TAPS):
        s = shift(u, dz, dy, dx)
        dW1[t] = s.reshape(CI, -1) @ dh.reshape(C, -1).T
    return loss, dW1, dW2, out
