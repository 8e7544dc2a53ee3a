"""INTEGRATION.md's Rust binding declares every entry point and constant the
public header declares, so the stub a maintainer copies cannot drift from
include/diaphane_amd.h."""
from __future__ import annotations

import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def header_text() -> str:
    return "\n".join(re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S) for h in (ROOT / "include").glob("*.h"))


def rust_block() -> str:
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```rust\n(.*?)```", text, flags=re.S)
    assert blocks, "INTEGRATION.md has no Rust binding"
    return "\n".join(blocks)


def test_every_entry_point_is_bound():
    declared = set(re.findall(r"\b(dp_[a-z0-9_]+)\s*\(", header_text()))
    bound = set(re.findall(r"pub fn (dp_[a-z0-9_]+)\(", rust_block()))
    assert declared, "no entry points found in the header"
    assert not declared - bound, f"missing from the Rust stub: {sorted(declared - bound)}"
    assert not bound - declared, f"not in the header: {sorted(bound - declared)}"


def test_every_constant_is_bound_with_its_value():
    values = dict(re.findall(r"\b(DP_[A-Z0-9_]+)\s*=\s*(-?\d+)", header_text()))
    bound = dict(re.findall(r"pub const (DP_[A-Z0-9_]+): i32 = (-?\d+);", rust_block()))
    assert values == bound
