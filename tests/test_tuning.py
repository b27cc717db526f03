"""The autotuner's candidate shapes are exactly the tile shapes compiled
into the tiled kernels, and every stored tuning entry names one of them
(a shape missing from the library would make the context fall back to the
untiled path without notice)."""
import json
import os
import re

from aperture3_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "aperture3_amd", "csrc", "kern_tile_mixed.hip")


def compiled_shapes():
    text = open(SRC).read()
    out = {}
    for dim, macro in ((2, "TILE2D_SHAPES"), (3, "TILE3D_SHAPES")):
        block = text[text.index(f"#define {macro}(X)"):]
        block = block[:block.index("\n\n")]
        shapes = set()
        for args in re.findall(r"X\(([^)]*)\)", block):
            v = [int(a) for a in args.split(",")]
            if dim == 2:
                shapes.add(((v[0], v[1], 0), v[2], v[3]))
            else:
                shapes.add(((v[0], v[0], v[0]), v[1], v[2]))
        out[dim] = shapes
    return out


def test_candidates_are_compiled():
    shapes = compiled_shapes()
    for dim in (2, 3):
        assert set(autotune.CANDIDATES[dim]) == shapes[dim], dim


def test_stored_entries_name_compiled_shapes():
    shapes = compiled_shapes()
    db = json.load(open(autotune.DEFAULT_DB))
    for k, e in db.items():
        if k.startswith("batch"):
            continue
        dim = int(k[1])
        assert (tuple(e["tile"]), e["halo"], e["threads"]) in shapes[dim], k
