"""Writes the golden fixtures under tests/golden/ from SoftGLRender's own
RendererSoft (the serial build compiled from the reference sources by
oracle/Makefile). Run in a container where /root/reference is present:

    python tests/golden/make_golden.py [name ...]   (default: every fixture)

Each fixture holds the outputs (RHS_READ order) of one seeded scene from
softglrender_amd/scenes.py; the scene is rebuilt from its name and parameters
when a test runs, so only outputs are stored.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from softglrender_amd import scenes  # noqa: E402
import oracle  # noqa: E402

# fixture name -> scene factory (keep in sync with tests/golden_scenes.py)
from tests.golden_scenes import GOLDEN  # noqa: E402


def main():
    if not oracle.have_reference():
        oracle.build(with_reference=True)
    names = sys.argv[1:] or list(GOLDEN)
    for name in names:
        make = GOLDEN[name]
        sc = make()
        out, _ = oracle.ref_replay(sc.stream.bytes(), sc.stream.output_bytes())
        arrays = sc.stream.split_outputs(out)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, *arrays)
        print(f"{name}: {[a.shape for a in arrays]} -> {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
