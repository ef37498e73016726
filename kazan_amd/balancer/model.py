"""cost_model.bin: the KzCostModel blob the driver loads at device creation.

Layout (little-endian, 19,616 bytes; device_common.h KzCostModel):
  u32 magic "KZCM", u32 version 1, u32 size, f32 frag_weight,
  f32 feat_mean[16], f32 feat_scale[16],
  bf16 w1[64][16], bf16 w2[64][64], bf16 w3[64][64]   (stored [out][in])
  f32 b1[64], b2[64], b3[64], w4[64], f32 b4, u32 pad[3]

The parameter vector of the training kernel (KAZAN_COST_PARAMS = 9,473
floats) is w1, w2, w3, b1, b2, b3, w4, b4 in that order, fp32.
"""
import numpy as np

FEATURES = 16
HIDDEN = 64
MAGIC = 0x4D435A4B
VERSION = 1
BLOB_BYTES = 19616

P_W1 = 0
P_W2 = P_W1 + HIDDEN * FEATURES
P_W3 = P_W2 + HIDDEN * HIDDEN
P_B1 = P_W3 + HIDDEN * HIDDEN
P_B2 = P_B1 + HIDDEN
P_B3 = P_B2 + HIDDEN
P_W4 = P_B3 + HIDDEN
P_B4 = P_W4 + HIDDEN
PARAMS = P_B4 + 1

FEATURE_NAMES = ["log2_1p_tris", "log2_1p_overlap_px", "hist_le2", "hist_le8", "hist_le32", "hist_le128",
                 "hist_gt128", "shared_frac", "z_range", "log2_samples_half", "blend", "depth_test",
                 "varyings_16", "textures_4", "ubos_4", "inside_frac"]


def to_bf16_bits(x):
    """f32 -> bf16 bit patterns, round to nearest even (the kernels' rule)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(x):
    """x rounded to bf16, as f32."""
    return from_bf16_bits(to_bf16_bits(x))


def init_params(seed=1):
    """He-initialised weights, zero biases (fp32 parameter vector)."""
    r = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[P_W1:P_W2] = r.normal(0, np.sqrt(2.0 / FEATURES), HIDDEN * FEATURES)
    p[P_W2:P_W3] = r.normal(0, np.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    p[P_W3:P_B1] = r.normal(0, np.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    p[P_W4:P_B4] = r.normal(0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    return p


def pack(params, feat_mean, feat_scale, frag_weight):
    """The blob for a trained parameter vector."""
    p = np.asarray(params, np.float32)
    out = bytearray()
    out += np.array([MAGIC, VERSION, BLOB_BYTES], np.uint32).tobytes()
    out += np.array([frag_weight], np.float32).tobytes()
    out += np.asarray(feat_mean, np.float32).tobytes() + np.asarray(feat_scale, np.float32).tobytes()
    out += to_bf16_bits(p[P_W1:P_B1]).tobytes()
    out += p[P_B1:P_B4].tobytes() + p[P_B4:P_B4 + 1].tobytes()
    out += np.zeros(3, np.uint32).tobytes()
    assert len(out) == BLOB_BYTES
    return bytes(out)


def unpack(blob):
    """dict of the blob's fields (weights as f32 values of their bf16 bits)."""
    assert len(blob) == BLOB_BYTES
    hdr = np.frombuffer(blob, np.uint32, 3, 0)
    if hdr[0] != MAGIC or hdr[1] != VERSION or hdr[2] != BLOB_BYTES:
        raise ValueError("not a version-1 KzCostModel blob")
    o = 12
    m = {"frag_weight": float(np.frombuffer(blob, np.float32, 1, o)[0])}
    o += 4
    m["feat_mean"] = np.frombuffer(blob, np.float32, FEATURES, o); o += 64
    m["feat_scale"] = np.frombuffer(blob, np.float32, FEATURES, o); o += 64
    for name, n, shape in (("w1", HIDDEN * FEATURES, (HIDDEN, FEATURES)), ("w2", HIDDEN * HIDDEN, (HIDDEN, HIDDEN)),
                           ("w3", HIDDEN * HIDDEN, (HIDDEN, HIDDEN))):
        m[name] = from_bf16_bits(np.frombuffer(blob, np.uint16, n, o)).reshape(shape)
        o += 2 * n
    for name in ("b1", "b2", "b3", "w4"):
        m[name] = np.frombuffer(blob, np.float32, HIDDEN, o)
        o += 4 * HIDDEN
    m["b4"] = float(np.frombuffer(blob, np.float32, 1, o)[0])
    return m


def predict_reference(blob, features):
    """numpy restatement of kz_cost_mlp: bf16 inputs and activations, fp32
    accumulation, ReLU; returns predicted fragments per row. Summation
    order differs from the MFMA's, so agreement is to rounding, not bits."""
    m = unpack(blob)
    x = bf16((np.asarray(features, np.float32) - m["feat_mean"]) * m["feat_scale"])
    h = x
    for w, b in (("w1", "b1"), ("w2", "b2"), ("w3", "b3")):
        h = bf16(np.maximum(h.astype(np.float64) @ m[w].T.astype(np.float64) + m[b], 0.0).astype(np.float32))
    y = h.astype(np.float64) @ m["w4"].astype(np.float64) + m["b4"]
    return np.maximum(np.exp2(y) - 1.0, 0.0)
