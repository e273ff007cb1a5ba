"""Writes assignment_targets.json: the expected node assignments of the reference's own
data-assignment test (tests/rustalgos/test_data.py:14-262, targets "visually confirmed in plots"
there), read from that file as text (ast.literal_eval of its module-level dicts; nothing is
imported or run). Keys are node indices, values the data keys assigned to that node.

Run from the repository root: python tests/golden/make_assignment_targets.py
"""

from __future__ import annotations

import ast
import json
from pathlib import Path

SRC = Path("/root/reference/tests/rustalgos/test_data.py")
NAMES = ("primal_targets", "decomp_targets", "decomp_targets_with_barriers", "targets_poly", "targets_poly_0")


def main() -> None:
    tree = ast.parse(SRC.read_text())
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name) and node.targets[0].id in NAMES:
            value = ast.literal_eval(node.value)
            out[node.targets[0].id] = {str(k): sorted(v) for k, v in value.items()}
    missing = set(NAMES) - set(out)
    if missing:
        raise SystemExit(f"not found: {sorted(missing)}")
    out["_source"] = "tests/rustalgos/test_data.py:14-262 of the reference (expected node assignments)"
    Path(__file__).with_name("assignment_targets.json").write_text(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
