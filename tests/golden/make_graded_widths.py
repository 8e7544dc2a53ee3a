"""Writes graded_256_widths.json: the oracle's restated `grade()` widths of the
256³ graded scattering grid, as f32 bit patterns, with the grid's recipe.

The reference ships no vector for these bits, and its Rust cannot be built
here; the fixture pins the restatement against drift, and the test beside it
asserts the invariants of the reference's own graded-mesh tests on it.

    python tests/golden/make_graded_widths.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from diaphane_amd import workloads  # noqa: E402


def main() -> None:
    grid = workloads.graded_scattering_256().scene.grid
    extent, widths, dt = oracle.grid_resolve(grid)
    spec = grid.spec()
    refinement = spec.refinements[0]
    out = {
        "about": "oracle grid_resolve of the graded_scattering_256 grid: primary widths x|y|z "
                 "and the time step, as f32 bit patterns (hex)",
        "recipe": {
            "base_extent": [int(v) for v in spec.extent],
            "cell_size": float(spec.cell_size),
            "courant": float(spec.courant),
            "max_ratio": float(spec.max_ratio),
            "refinement": {"center": [float(v) for v in refinement.center],
                           "size": [float(v) for v in refinement.size],
                           "cell_size": float(refinement.cell_size)},
        },
        "extent": list(extent),
        "widths_f32_hex": [f"{v:08x}" for v in widths.view(np.uint32).tolist()],
        "dt_f32_hex": f"{int(np.float32(dt).view(np.uint32)):08x}",
    }
    path = Path(__file__).with_name("graded_256_widths.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path} ({len(widths)} widths)")


if __name__ == "__main__":
    main()
