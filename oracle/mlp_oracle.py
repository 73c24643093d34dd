"""CPU oracle of the learned-start MLP (TEST INFRASTRUCTURE ONLY; the product never imports it).

Restates the network of arpeggio_amd/learned.py in numpy: bf16 rounding (round to nearest even of the float32 bit
pattern) at the same points as the GPU (inputs, weights, GELU outputs, output-layer and hidden gradients), float64
arithmetic elsewhere, GELU in tanh form, mean binary cross-entropy against +-1 spins. The learned start has no
reference counterpart; this oracle is pinned by central finite differences of its own float64 loss (tests)."""
import numpy as np

C_GELU = 0.7978845608028654


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * x * x)


def forward(p, X, round_bf16=True):
    r = bf16 if round_bf16 else (lambda v: v)
    W1, W2, W3 = (r(p[k]).astype(np.float64) for k in ('W1', 'W2', 'W3'))
    X = r(X).astype(np.float64)
    A1 = X @ W1 + p['b1']
    H1 = r(gelu(A1).astype(np.float32)).astype(np.float64) if round_bf16 else gelu(A1)
    A2 = H1 @ W2 + p['b2']
    H2 = r(gelu(A2).astype(np.float32)).astype(np.float64) if round_bf16 else gelu(A2)
    Z = H2 @ W3 + p['b3']
    return Z, dict(X=X, A1=A1, H1=H1, A2=A2, H2=H2, W1=W1, W2=W2, W3=W3)


def loss_and_grads(p, X, Y, round_bf16=True):
    r = (lambda v: bf16(v.astype(np.float32)).astype(np.float64)) if round_bf16 else (lambda v: v)
    Z, S = forward(p, X, round_bf16)
    t = (np.asarray(Y, dtype=np.float64) + 1) / 2
    n = Z.size
    loss = float((np.logaddexp(0, Z) - t * Z).sum() / n)
    dZf = (1 / (1 + np.exp(-Z)) - t) / n
    dZ = r(dZf)
    g = dict(W3=S['H2'].T @ dZ, b3=dZf.sum(0))
    dH2 = dZ @ S['W3'].T
    dA2f = dH2 * gelu_grad(S['A2'])
    dA2 = r(dA2f)
    g['W2'] = S['H1'].T @ dA2
    g['b2'] = dA2f.sum(0)
    dH1 = dA2 @ S['W2'].T
    dA1f = dH1 * gelu_grad(S['A1'])
    dA1 = r(dA1f)
    g['W1'] = S['X'].T @ dA1
    g['b1'] = dA1f.sum(0)
    return loss, g


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v
