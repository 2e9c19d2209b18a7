"""CPU reference of the scheduler's MLP cost model (numpy).

`forward_backward(params, tapes, targets, bf16=True)` rounds to bf16 at the
same points as the HIP kernels (features, weights E/W1/W2, activations p, a1,
a2 and the backward signals d2, d1, dp); bf16=False computes the same
function in float64 from the bf16-rounded inputs and weights.
"""
import numpy as np

IN, HID = 256, 256
OFF_E = 0
OFF_W1 = OFF_E + IN * HID
OFF_B1 = OFF_W1 + HID * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NUM_PARAMS = OFF_B3 + 1


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def features(tapes):
    """Byte histogram / 128, shape (n, 256)."""
    t = np.asarray(tapes, dtype=np.int64).reshape(-1, 128)
    n = t.shape[0]
    x = np.zeros((n, IN), dtype=np.float64)
    np.add.at(x, (np.repeat(np.arange(n), 128), t.reshape(-1)), 1.0)
    return (x / 128).astype(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return dict(E=p[OFF_E:OFF_W1].reshape(IN, HID), W1=p[OFF_W1:OFF_B1].reshape(HID, HID),
                b1=p[OFF_B1:OFF_W2], W2=p[OFF_W2:OFF_B2].reshape(HID, HID),
                b2=p[OFF_B2:OFF_W3], w3=p[OFF_W3:OFF_B3], b3=p[OFF_B3:NUM_PARAMS])


def forward_backward(params, tapes, targets, bf16_points=True, bf16_weights=True):
    """bf16_weights=False keeps E, W1, W2 unrounded (a smooth function of the
    parameters, for finite differences)."""
    P = unpack(params)
    r = bf16 if bf16_points else (lambda v: np.asarray(v, dtype=np.float64))
    f = np.float32 if bf16_points else np.float64
    x = features(tapes).astype(f)
    qw = bf16 if bf16_weights else (lambda v: v)
    E, W1, W2 = (qw(P[k]).astype(f) for k in ("E", "W1", "W2"))
    b1, b2, w3 = (P[k].astype(f) for k in ("b1", "b2", "w3"))
    b3 = f(P["b3"][0])
    p = r(x @ E)
    a1 = r(np.maximum(p @ W1 + b1, 0))
    a2 = r(np.maximum(a1 @ W2 + b2, 0))
    y = a2 @ w3 + b3
    n = x.shape[0]
    e = y - np.asarray(targets, dtype=f)
    loss = 0.5 * float(np.mean(e.astype(np.float64) ** 2))
    dy = e / n
    d2 = r(np.where(a2 != 0, dy[:, None] * w3[None, :], 0))
    g = {}
    g["w3"] = a2.T @ dy
    g["b3"] = np.array([dy.sum()])
    g["W2"] = a1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = r(np.where(a1 != 0, d2 @ W2.T, 0))
    g["W1"] = p.T @ d1
    g["b1"] = d1.sum(0)
    dp = r(d1 @ W1.T)
    g["E"] = x.T @ dp
    flat = np.concatenate([g["E"].ravel(), g["W1"].ravel(), g["b1"], g["W2"].ravel(),
                           g["b2"], g["w3"], g["b3"]]).astype(np.float64)
    return y.astype(np.float64), loss, flat


BLOCKS = {"E": (OFF_E, OFF_W1), "W1": (OFF_W1, OFF_B1), "b1": (OFF_B1, OFF_W2),
          "W2": (OFF_W2, OFF_B2), "b2": (OFF_B2, OFF_W3), "w3": (OFF_W3, OFF_B3),
          "b3": (OFF_B3, NUM_PARAMS)}


def predict_folded(params, tapes):
    """The scheduler's fused inference: first layer folded in fp32 and rounded
    to bf16 as one matrix, activations a1, a2 rounded to bf16."""
    P = unpack(params)
    x = features(tapes).astype(np.float32)
    EW1 = bf16((P["E"].astype(np.float64) @ P["W1"].astype(np.float64)).astype(np.float32))
    a1 = bf16(np.maximum(x @ EW1 + P["b1"], 0))
    a2 = bf16(np.maximum(a1 @ bf16(P["W2"]) + P["b2"], 0))
    return a2 @ P["w3"] + P["b3"][0]
