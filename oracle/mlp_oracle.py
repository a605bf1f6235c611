"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

numpy restatement of the learned parts of the Q1 preconditioner: the patch features, the network, its node
weights, the sub-cell scales of the embedded-Q1 matrix (network kind 1: 11 inputs, 3 outputs) and the gradient
of a squared-error loss of the raw outputs.  The reference has no learned component; these functions define the arithmetic the
HIP kernels must reproduce, and the tests pin them by finite differences (gradient) and by
closed forms (features, bf16 rounding, zero network).
"""
import numpy as np

NIN, NH, NOUT = 32, 64, 8
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT
CLAMP = 4.0


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    u = np.where(finite, u + 0x7FFF + ((u >> 16) & 1), u) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


# sizes per network kind: 0 = patch node weights, 1 = sub-cell scales of the embedded-Q1 matrix
SHAPES = {0: (NIN, NOUT), 1: (11, 3)}


def nparam(kind=0):
    nin, nout = SHAPES[kind]
    return NH * nin + NH + NH * NH + NH + nout * NH + nout


def unpack(w, kind=0):
    nin, nout = SHAPES[kind]
    w = np.asarray(w, dtype=np.float64)
    o1 = NH * nin
    o2 = o1 + NH
    o3 = o2 + NH * NH
    o4 = o3 + NH
    o5 = o4 + nout * NH
    assert w.size == o5 + nout
    return (w[:o1].reshape(NH, nin), w[o1:o2], w[o2:o3].reshape(NH, NH), w[o3:o4], w[o4:o5].reshape(nout, NH), w[o5:])


def silu(x):
    return x / (1 + np.exp(-x))


def dsilu(x):
    s = 1 / (1 + np.exp(-x))
    return s * (1 + x * (1 - s))


def forward(w, X, kind=0):
    W1, b1, W2, b2, W3, b3 = unpack(w, kind)
    Z1 = X @ W1.T + b1
    Z2 = silu(Z1) @ W2.T + b2
    return Z1, Z2, silu(Z2) @ W3.T + b3


def loss_and_grad(w, X, T, kind=0):
    """Mean squared error of the raw outputs over all B x NOUT entries and its gradient (float64)."""
    X = np.asarray(X, dtype=np.float64)
    B = X.shape[0]
    nout = SHAPES[kind][1]
    W1, b1, W2, b2, W3, b3 = unpack(w, kind)
    Z1, Z2, O = forward(w, X, kind)
    D = O - T
    loss = np.mean(D ** 2)
    dO = 2 * D / (B * nout)
    H1, H2 = silu(Z1), silu(Z2)
    dZ2 = (dO @ W3) * dsilu(Z2)
    dZ1 = (dZ2 @ W2) * dsilu(Z1)
    g = np.concatenate([(dZ1.T @ X).ravel(), dZ1.sum(0), (dZ2.T @ H1).ravel(), dZ2.sum(0), (dO.T @ H2).ravel(), dO.sum(0)])
    return loss, g


def patches(box):
    nx, ny, nz = box
    P, Q, R = np.meshgrid(np.arange(nx // 2), np.arange(ny // 2), np.arange(nz // 2), indexing="ij")
    P, Q, R = P.ravel(), Q.ravel(), R.ravel()
    m = np.arange(8)
    I = 2 * P[:, None] + (m >> 2)
    J = 2 * Q[:, None] + ((m >> 1) & 1)
    K = 2 * R[:, None] + (m & 1)
    return I, J, K, (I * ny + J) * nz + K  # each [npatch][8]


def features(box, p, lo, dinv):
    """[npatch][32]: log D_m - mean log D, then (lattice index mod p) / p per node and direction."""
    I, J, K, idx = patches(box)
    ld = -np.log(dinv[idx])
    pos = np.stack([((lo[0] + I) % p) / p, ((lo[1] + J) % p) / p, ((lo[2] + K) % p) / p], axis=2)  # [np][8][3]
    return np.concatenate([ld - ld.mean(1, keepdims=True), pos.reshape(len(idx), 24)], axis=1)


def node_weights(box, w, X):
    """s over the box nodes: exp(clamp(o)/2) on patch nodes, 1 elsewhere (bf16 weights)."""
    _, _, _, idx = patches(box)
    _, _, O = forward(bf16(w), X)
    s = np.ones(box[0] * box[1] * box[2])
    s[idx] = np.exp(0.5 * np.clip(O, -CLAMP, CLAMP))
    return s


def cell_scales(w, X):
    """Sub-cell derivative scales sigma = exp(clamp(o)) of a kind-1 network (bf16 weights) at features X."""
    _, _, O = forward(bf16(w), X, kind=1)
    return np.exp(np.clip(O, -CLAMP, CLAMP))
