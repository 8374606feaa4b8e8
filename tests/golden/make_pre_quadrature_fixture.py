"""Write ``pre_quadrature_reference_synthetic.json``: the two synthetic cases of the reference's whole-view
regression file (``tests/data/pre_quadrature_reference.json`` of the reference, read as JSON data; nothing of the
reference is imported or run). The real-library cases of that file are left out: they need the XCOM photon data.

Usage: ``python tests/golden/make_pre_quadrature_fixture.py [path to the reference checkout]``.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main() -> None:
    root = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    data = json.loads((root / "tests" / "data" / "pre_quadrature_reference.json").read_text())
    cases = {k: v for k, v in data["cases"].items() if k.startswith("synthetic_")}
    out = {"description": data["description"], "source": "tests/data/pre_quadrature_reference.json (reference)",
           "cases": cases}
    (HERE / "pre_quadrature_reference_synthetic.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
