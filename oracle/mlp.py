"""CPU restatement (float64 numpy) of the library's dense network: forward,
masked mean-squared-error loss, backward and Adam.

TEST INFRASTRUCTURE ONLY (the checker of csrc/hip/mlp.hip): used by tests/
and nothing else.  The reference has no network (its kernel choice is the
fixed rule of src/alto_dev.cu:1975,1996-2001), so this follows the
library's own definition:

  layer l: Y = X W_l^T + b_l, W_l row-major (w[l+1] x w[l]); ReLU after every
           layer but the last;
  params:  flat, layer by layer, W_l then b_l;
  loss:    sum(M (Y - T)^2) / sum(M) (sum(M) = 0 -> divide by 1);
  Adam:    m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;
           p -= lr sqrt(1-b2^t) / (1-b1^t) * m / (sqrt(v) + eps)
           (Kingma & Ba 2015, Algorithm 1, with the bias corrections folded
           into the step size as in their section 2).

Pinned by tests/test_mlp_oracle.py: central finite differences of the loss
against the analytic gradient for every parameter of small networks; a
one-layer linear network whose gradient has the closed form
2/sum(M) ((Y-T) M)^T [X 1]; Adam's first step equals -lr * sign(g) (to
eps) for every parameter.
"""
import numpy as np


def split(widths, p):
    """Views (W_l, b_l) of a flat parameter vector."""
    out, k = [], 0
    for l in range(len(widths) - 1):
        nin, nout = widths[l], widths[l + 1]
        W = p[k:k + nin * nout].reshape(nout, nin)
        k += nin * nout
        b = p[k:k + nout]
        k += nout
        out.append((W, b))
    assert k == len(p)
    return out


def nparams(widths):
    return sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(len(widths) - 1))


def forward(widths, p, X):
    """Returns the list of activations [X, a_1, ..., Y]."""
    acts = [np.asarray(X, np.float64)]
    layers = split(widths, np.asarray(p, np.float64))
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ W.T + b
        if l + 1 < len(layers):
            z = np.maximum(z, 0.0)
        acts.append(z)
    return acts


def loss(widths, p, X, T, M):
    Y = forward(widths, p, X)[-1]
    den = M.sum() if M.sum() > 0 else 1.0
    return float((M * (Y - T) ** 2).sum() / den)


def loss_grad(widths, p, X, T, M):
    """(loss, flat gradient)."""
    p = np.asarray(p, np.float64)
    acts = forward(widths, p, X)
    layers = split(widths, p)
    Y = acts[-1]
    den = M.sum() if M.sum() > 0 else 1.0
    L = float((M * (Y - T) ** 2).sum() / den)
    dY = 2.0 * M * (Y - T) / den
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads.append((dY.T @ acts[l], dY.sum(axis=0)))
        if l > 0:
            dX = dY @ W
            dY = dX * (acts[l] > 0)
    g = []
    for dW, db in reversed(grads):
        g.append(dW.ravel())
        g.append(db)
    return L, np.concatenate(g)


class Adam:
    def __init__(self, n, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.b1, self.b2, self.eps = b1, b2, eps

    def step(self, p, g, lr):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        lr_t = lr * np.sqrt(1 - self.b2 ** self.t) / (1 - self.b1 ** self.t)
        return p - lr_t * self.m / (np.sqrt(self.v) + self.eps)
