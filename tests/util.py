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


def assert_parity(got, want, label="", colour_frac=1e-4):
    """The parity bar: depth bit-exact, colour within 1/255 per channel on every
    value and exact on all but `colour_frac` of them."""
    dmis, cmax, cfrac = compare(got, want, label)
    assert dmis == 0, f"{label}: {dmis} depth values differ"
    assert cmax <= 1, f"{label}: colour differs by up to {cmax}"
    assert cfrac <= colour_frac, f"{label}: {cfrac:.2e} of colour values differ"


def load_golden(path):
    with np.load(path, allow_pickle=False) as z:
        return [z[f"arr_{i}"] for i in range(len(z.files))]
