"""The engine's multiplier encoders against PyTorch's conversions
(tests/golden/quant_vectors.json, written by make_quant_vectors.py beside it from
torch.bfloat16 and torch.float8_e4m3fn): BF16 round to nearest even, and FP8-e4m3
with the projection's power-of-two storage scale, bit for bit; the row-sum factor
is the generated over the stored multiplier sum."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "quant_vectors.json")


def floats(hexes):
    return np.array([int(h, 16) for h in hexes], np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def groups():
    with open(GOLDEN) as f:
        return json.load(f)["groups"]


def test_bf16_matches_torch(pkg, groups):
    for g in groups:
        x = floats(g["input"])
        got, scale, ss = pkg.encode_multipliers(pkg.WEIGHTS_BF16, x)
        want = floats(g["bf16"])
        bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
        assert bad.size == 0, f"{g['name']}: {x[bad[:5]]} -> {got[bad[:5]]}, torch {want[bad[:5]]}"
        assert scale == 1.0
        assert np.float32(ss) == np.float32(x.astype(np.float64).sum() / want.astype(np.float64).sum())


def test_fp8_e4m3_matches_torch(pkg, groups):
    for g in groups:
        x = floats(g["input"])
        got, scale, ss = pkg.encode_multipliers(pkg.WEIGHTS_FP8, x)
        assert scale == g["fp8_scale"], g["name"]
        # a power of two that brings the largest magnitude to (224, 448]
        assert np.log2(scale) == np.round(np.log2(scale)) and 224 < np.abs(x).max() / scale <= 448
        want = floats(g["fp8"])
        bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
        assert bad.size == 0, f"{g['name']}: {x[bad[:5]]} -> {got[bad[:5]]}, torch {want[bad[:5]]}"
        assert np.float32(ss) == np.float32(x.astype(np.float64).sum() / want.astype(np.float64).sum())


def test_f32_is_stored_as_given(pkg, groups):
    x = floats(groups[0]["input"])
    got, scale, ss = pkg.encode_multipliers(pkg.WEIGHTS_F32, x)
    assert np.array_equal(got.view(np.uint32), x.view(np.uint32)) and scale == 1.0 and ss == 1.0


def test_fp8_saturates_instead_of_nan(pkg):
    """Beyond 448 x scale torch's e4m3fn conversion gives NaN; the encoder only
    ever sees values within 448 x scale (the scale is chosen from the largest
    magnitude), and for a lone outlier past that bound it saturates."""
    x = np.array([1.0, 2.0, 448.0], np.float32)
    got, scale, _ = pkg.encode_multipliers(pkg.WEIGHTS_FP8, x)
    assert scale == 1.0 and np.array_equal(got, x)
    with pytest.raises(pkg.ScaffoldError):
        pkg.encode_multipliers(pkg.WEIGHTS_FP8, np.array([1.0, np.nan], np.float32))
