"""Writes tests/golden/builder_vectors.npz: input/output vectors of the
reference's StringBuilder (mepgl/mepgl_lib/builder.py), for pinning
mepgl_amd.builder.  Run in the build container only, where the reference
tree exists; the vectors (data only) are committed, the reference is not.

    python tests/golden/make_builder_vectors.py [/path/to/reference]

Each case fixes numpy's global seed before generate_matter_fields, as a
seeded config.py would; the same case description drives the test.
"""
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from builder_cases import CASES, build  # noqa: E402


def main(argv):
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    path = os.path.join(ref, "mepgl", "mepgl_lib", "builder.py")
    spec = importlib.util.spec_from_file_location("reference_builder", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {"cases": np.array(json.dumps(CASES))}
    for k, case in enumerate(CASES):
        for name, arr in build(mod.StringBuilder, case).items():
            out[f"c{k}_{name}"] = arr
    dst = os.path.join(HERE, "builder_vectors.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv)
