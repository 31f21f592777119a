"""CPU restatement of the pivot scorer's network and its gradient -- TEST
INFRASTRUCTURE ONLY (the product trains and runs the network in HIP kernels).

The scorer has no reference counterpart (the reference's pivot is an atomicCAS
race, scc_routines.cu:214-221); this module is the checker for the build's own
kernels: MLP 12 -> 32 -> 32 -> 1, ReLU, logit output, loss = mean binary
cross-entropy of sigmoid(logit).  With ``bf16=True`` weights and hidden
activations are rounded to bfloat16 (round to nearest even) exactly where the
HIP kernels round them; the arithmetic is float64.

Pinning: ``grad`` (bf16=False) is checked against central finite differences
of ``loss`` in tests/test_oracle_mlp.py; ``bf16_round`` against hand-worked
bit patterns.
"""
import numpy as np

NF, NH = 12, 32
SIZES = [("W1", (NH, NF)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, X, bf16=True):
    P = unpack(bf16_round(p) if bf16 else p)
    X = np.asarray(X, np.float64)
    a1 = X @ P["W1"].T + P["b1"]
    h1 = np.maximum(a1, 0.0)
    if bf16:
        h1 = bf16_round(h1)
    a2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(a2, 0.0)
    if bf16:
        h2 = bf16_round(h2)
    s = h2 @ P["w3"] + P["b3"][0]
    return s, h1, h2, P


def loss(p, X, y, bf16=False):
    s, _, _, _ = forward(p, X, bf16)
    y = np.asarray(y, np.float64)
    return float(np.mean(np.maximum(s, 0) - s * y + np.log1p(np.exp(-np.abs(s)))))


def grad(p, X, y, bf16=False):
    """Analytic gradient of the mean BCE; the same backward pass as k_mlp_fwdbwd/k_mlp_grad."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    N = len(X)
    s, h1, h2, P = forward(p, X, bf16)
    ds = (1.0 / (1.0 + np.exp(-s)) - y) / N
    dh2 = np.outer(ds, P["w3"]) * (h2 > 0)
    dh1 = (dh2 @ P["W2"]) * (h1 > 0)
    g = {"W1": dh1.T @ X, "b1": dh1.sum(0), "W2": dh2.T @ h1, "b2": dh2.sum(0),
         "w3": h2.T @ ds, "b3": np.array([ds.sum()])}
    return pack(g)


def init_params(seed=0):
    """He-style initialisation (the trainer's own, restated for tests)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / NF), (NH, NF)), "b1": np.zeros(NH),
         "W2": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b2": np.zeros(NH),
         "w3": rng.normal(0, np.sqrt(1.0 / NH), NH), "b3": np.zeros(1)}
    return pack(d).astype(np.float32)
