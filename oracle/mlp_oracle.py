"""CPU restatement of the surrogate MLP (test infrastructure only).

The surrogate has no counterpart in the reference: this file states what the
HIP kernels of dmipy-sim_amd/csrc/mlp.hip compute, in float64 numpy, with the
bfloat16 roundings placed where the kernels place them:

  forward   Xb = bf(x);  Z_l = Y_l . bf(W_l) + b_l;  Y_{l+1} = bf(gelu(Z_l));
            y = Y_D . w_out + b_out
  backward  dy = 2 (y - t) / B  (mean squared error)
            dw_out = Y_D^T dy, db_out = sum dy
            dZ_{D-1} = bf(dy w_out^T * gelu'(Z_{D-1}))
            dW_l = Y_l^T dZ_l,  db_l = colsum dZ_l
            dZ_{l-1} = bf((dZ_l . bf(W_l)^T) * gelu'(Z_{l-1}))
  Adam      m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;
            p -= lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps)

GELU is the tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
With ``bf16=False`` every rounding is skipped; the CPU tests pin that exact
gradient against central finite differences of the forward pass.

Public parameter layout (float32, flat): for each hidden layer W_l (K_l x
width, row-major, K_0 = n_in) then b_l (width); then w_out (width), b_out (1).
"""
import numpy as np

_C = 0.7978845608028654  # sqrt(2/pi)


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    th = np.tanh(_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * _C * (1.0 + 3 * 0.044715 * x * x)


def num_params(n_in, width, depth):
    n = n_in * width + width
    n += (depth - 1) * (width * width + width)
    return n + width + 1


def unpack(params, n_in, width, depth):
    """Split the flat public vector into ([W_l], [b_l], w_out, b_out)."""
    p = np.asarray(params, dtype=np.float64)
    Ws, bs, off = [], [], 0
    for l in range(depth):
        K = n_in if l == 0 else width
        Ws.append(p[off:off + K * width].reshape(K, width))
        off += K * width
        bs.append(p[off:off + width])
        off += width
    w_out = p[off:off + width]
    b_out = p[off + width]
    assert off + width + 1 == p.size
    return Ws, bs, w_out, b_out


def pack(Ws, bs, w_out, b_out):
    parts = []
    for W, b in zip(Ws, bs):
        parts += [np.ravel(W), np.ravel(b)]
    parts += [np.ravel(w_out), np.atleast_1d(b_out)]
    return np.concatenate(parts)


def forward(params, x, n_in, width, depth, bf16=True, keep=False):
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Ws, bs, w_out, b_out = unpack(params, n_in, width, depth)
    Y = [rnd(x)]
    Z = []
    for l in range(depth):
        z = Y[-1] @ rnd(Ws[l]) + bs[l]
        Z.append(z)
        Y.append(rnd(gelu(z)))
    y = Y[-1] @ w_out + b_out
    return (y, Y, Z) if keep else y


def gradients(params, x, t, n_in, width, depth, bf16=True):
    """(grads in the public layout, mean squared error)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Ws, bs, w_out, b_out = unpack(params, n_in, width, depth)
    y, Y, Z = forward(params, x, n_in, width, depth, bf16=bf16, keep=True)
    B = x.shape[0]
    e = y - np.asarray(t, np.float64)
    dy = 2.0 * e / B
    g_wout = Y[depth].T @ dy
    g_bout = dy.sum()
    dZ = rnd(dy[:, None] * w_out[None, :] * gelu_grad(Z[depth - 1]))
    gW = [None] * depth
    gb = [None] * depth
    for l in range(depth - 1, -1, -1):
        gW[l] = Y[l].T @ dZ
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dZ = rnd((dZ @ rnd(Ws[l]).T) * gelu_grad(Z[l - 1]))
    return pack(gW, gb, g_wout, g_bout), float(np.mean(e * e))


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (t counts from 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
