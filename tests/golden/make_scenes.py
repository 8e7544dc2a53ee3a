"""Copies the reference's curated scene files (scenes/*.ron, the data its
tests/scenes.rs loads) into tests/golden/scenes/ as fixtures.

Run in a checkout that has the reference next to it:

    python tests/golden/make_scenes.py /root/reference
"""
from __future__ import annotations

import shutil
import sys
from pathlib import Path


def main() -> None:
    source = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "scenes"
    target = Path(__file__).with_name("scenes")
    target.mkdir(exist_ok=True)
    files = sorted(source.glob("*.ron"))
    for path in files:
        shutil.copyfile(path, target / path.name)
    print(f"copied {len(files)} scene files to {target}")


if __name__ == "__main__":
    main()
