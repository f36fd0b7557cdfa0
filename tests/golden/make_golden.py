"""Writes the committed fixtures under tests/golden/.

  pcg32.json        PCG32 streams for (pixel, sample, seed) keys, computed by
                    the Python restatement in tests/test_oracle_math.py
                    (independent of the C header)
  karras8.json      the 8-key example of Karras 2012 ("Maximizing
                    Parallelism in the Construction of BVHs, Octrees, and k-d
                    Trees", Fig. 3), children derived by hand
  cornell_lbvh.npz  Morton codes, sorted ids and tree of the 36-triangle
                    Cornell box, from the CPU oracle
  trace_soup.npz    2048 closest hits on a seeded 4096-triangle soup, from
                    the CPU oracle
  cornell_64.npy    64x64, 4 spp, depth 8 Cornell accumulation (FP32), from
                    the CPU oracle

The reference ships no tests or vectors (it has no test directory), so only
the first two are independent of this build; the rest freeze the CPU
oracle's output so that a later change to it shows up as a diff.

Run from the repository root: python tests/golden/make_golden.py
"""
import importlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    from test_oracle_math import MASK64, pcg32_python
    import pyoracle

    mcpt = importlib.import_module("monocuda-pt_amd")
    cases = []
    for pixel, sample, seed in [(0, 0, 0), (65535, 15, 0), (777, 3, 1234), (1048575, 63, 1)]:
        cases.append({"pixel": pixel, "sample": sample, "seed": seed,
                      "out": pcg32_python(((sample << 32) ^ seed) & MASK64, pixel, 16)})
    with open(os.path.join(HERE, "pcg32.json"), "w") as f:
        json.dump({"cases": cases}, f, indent=1)

    karras = {
        "keys": [0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110],
        # internal node i: (left, left_is_leaf, right, right_is_leaf)
        "nodes": [[3, 0, 4, 0], [0, 1, 1, 1], [2, 1, 3, 1], [1, 0, 2, 0], [4, 1, 5, 0], [6, 0, 7, 1], [5, 1, 6, 1]],
        "parent": [-1, 3, 3, 0, 0, 4, 5],
    }
    with open(os.path.join(HERE, "karras8.json"), "w") as f:
        json.dump(karras, f, indent=1)

    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    s = pyoracle.Scene(hot, cold, mats)
    d = s.dump()
    np.savez_compressed(os.path.join(HERE, "cornell_lbvh.npz"),
                        **{k: d[k] for k in ("morton", "sorted_keys", "sorted_ids", "left", "right", "left_is_leaf",
                                             "right_is_leaf", "parent", "leaf_parent")})
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, 64, 64)
    img, _, _ = s.render(pyoracle.camera_init(desc), 4, 8, seed=3)
    np.save(os.path.join(HERE, "cornell_64.npy"), img)

    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=5, param=4096)
    s = pyoracle.Scene(hot, cold, mats)
    rng = np.random.default_rng(5)
    rays = np.zeros(2048, pyoracle.RAY_DTYPE)
    rays["origin"] = rng.uniform(-0.1, 1.1, (2048, 3))
    dd = rng.normal(size=(2048, 3))
    rays["dir"] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    rays["tmax"] = 1e30
    hits, steps, leaves = s.trace_closest(rays)
    np.savez_compressed(os.path.join(HERE, "trace_soup.npz"), rays=rays, hits=hits, steps=steps, leaves=leaves)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
