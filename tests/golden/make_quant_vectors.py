"""Writes quant_vectors.json: expected BF16 and FP8-e4m3 storage of synaptic
multipliers, computed by PyTorch's own conversions (torch.bfloat16 and
torch.float8_e4m3fn, round to nearest even), to pin the engine's encoders
(scaffold_encode_multipliers, the planner's encoding step).

Inputs: the first multipliers the seeded scaffold generator draws (U(0.8, 1.2),
the 9.9k-cell scaffold's parallel-fibre-to-Purkinje projection), the same
scaled into other binades, and edge cases: every midpoint between adjacent
representable values of each format near 1 (ties to even), BF16 values near the
bottom of the float range, e4m3 subnormals and their midpoints, and values up to
the largest e4m3 value 448.  FP8 groups carry their power-of-two scale
2^ceil(log2(max|x| / 448)); torch converts x / scale, and the stored multiplier
is that value times the scale.

Run in the build container (torch and the built library present):
    python tests/golden/make_quant_vectors.py
"""
import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def hexs(a):
    return [format(int(u), "08x") for u in np.ascontiguousarray(a, np.float32).view(np.uint32)]


def torch_bf16(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).float().numpy()


def torch_fp8(x, scale):
    y = torch.from_numpy(np.ascontiguousarray(x / np.float32(scale), np.float32)).to(torch.float8_e4m3fn).float()
    return (y.numpy() * np.float32(scale)).astype(np.float32)


def fp8_scale(x):
    mx = float(np.abs(x).max())
    return float(2.0 ** np.ceil(np.log2(mx / 448.0))) if mx > 0 else 1.0


def main():
    import __graft_entry__ as ge
    pkg = ge.load_package()
    env = pkg.NetworkEnvironment(1, 7, 16, 2, 0)
    env.describe_synthetic(pkg.synth_preset(0))
    q = next(i for i in range(env.info().n_projections) if env.projection(i).type == 0)
    gen = env.projection_csr(q)[2][:600].astype(np.float32)
    env.finalize()
    assert gen.min() >= 0.8 and gen.max() <= 1.2

    groups = []
    # generated multipliers, and the same in other binades
    for k, mult in enumerate((1.0, 0.02, 75.0, 2.0 ** -20)):
        groups.append(("generated x %g" % mult, (gen * np.float32(mult)).astype(np.float32)))
    # BF16 ties: midpoints between consecutive bf16 values in [1, 2) and [0.5, 1)
    m = np.arange(0, 128, dtype=np.uint32)
    lo = ((0x3F80 + m) << 16).astype(np.uint32)
    ties = (lo + 0x8000).view(np.float32)
    lo2 = ((0x3F00 + m) << 16).astype(np.uint32)
    ties2 = (lo2 + 0x8000).view(np.float32)
    near = (lo + np.uint32(0x7FFF)).view(np.float32)           # just below a midpoint
    groups.append(("bf16 midpoints and neighbours", np.concatenate([ties, ties2, near])))
    # e4m3 codes: every positive finite value, their midpoints, subnormals (scale 1)
    codes = []
    for c in range(1, 0x7F):
        e, mm = (c >> 3) & 15, c & 7
        codes.append((1 + mm / 8) * 2.0 ** (e - 7) if e else mm * 2.0 ** -9)
    codes = np.array(codes, np.float64)
    mids = (codes[:-1] + codes[1:]) / 2
    sub = np.array([2.0 ** -10, 3 * 2.0 ** -11, 2.0 ** -12, 5 * 2.0 ** -12], np.float64)
    e4 = np.concatenate([codes, mids, sub, [448.0]]).astype(np.float32)
    groups.append(("e4m3 values, midpoints, subnormals", e4))

    out = {"note": "float32 bit patterns (hex); expected values from torch.bfloat16 / torch.float8_e4m3fn",
           "torch": torch.__version__, "groups": []}
    for name, x in groups:
        x = np.ascontiguousarray(x, np.float32)
        s = fp8_scale(x)
        out["groups"].append({"name": name, "input": hexs(x), "bf16": hexs(torch_bf16(x)),
                              "fp8_scale": s, "fp8": hexs(torch_fp8(x, s))})
    with open(os.path.join(HERE, "quant_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", sum(len(g["input"]) for g in out["groups"]), "values")


if __name__ == "__main__":
    main()
