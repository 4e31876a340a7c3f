"""Comparison helpers shared by the parity tests."""
import numpy as np


def compare(got, want, label=""):
    """Returns (depth_mismatches, colour_max_diff, colour_diff_fraction) over a list of outputs."""
    dmis, cmax, cfrac = 0, 0, 0.0
    for g, w in zip(got, want):
        assert g.shape == w.shape, (label, g.shape, w.shape)
        if g.dtype == np.float32:
            dmis += int((g.view(np.uint32) != w.view(np.uint32)).sum())
        else:
            d = np.abs(g.astype(np.int16) - w.astype(np.int16))
            cmax = max(cmax, int(d.max()))
            cfrac = max(cfrac, float((d != 0).mean()))
    return dmis, cmax, cfrac


def differing_values(got, want):
    """Colour values (bytes of RGBA8 outputs) that differ."""
    return sum(int((g != w).sum()) for g, w in zip(got, want) if g.dtype == np.uint8)


def assert_identical(got, want, label=""):
    """Every output identical bit for bit (depth, colour, samples)."""
    assert len(got) == len(want), label
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, (label, i, g.shape, w.shape)
        bad = int((g.view(np.uint8) != w.view(np.uint8)).sum())
        assert bad == 0, f"{label}: output {i} differs in {bad} bytes"


def load_golden(path):
    with np.load(path, allow_pickle=False) as z:
        return [z[f"arr_{i}"] for i in range(len(z.files))]
