"""Write tests/golden/dirbe_25um_bandpass.json: the DIRBE 25 um band's bandpass.

The reference's test data hold the 25 um DIRBE bandpass as 82 wavelength/weight
pairs (tests/dirbe_tabulated.py:158-328, ``DIRBE_25um_BANDPASS``; no test of the
reference uses it). The file is read as text with ``ast`` -- nothing of the
reference is imported or executed -- and the two number lists are written out.

    python tests/golden/make_dirbe_bandpass.py [/root/reference/tests/dirbe_tabulated.py]
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

REFERENCE_TABLE = Path("/root/reference/tests/dirbe_tabulated.py")
OUT = Path(__file__).resolve().parent / "dirbe_25um_bandpass.json"


def bandpass(path: Path) -> tuple[list[float], list[float]]:
    tree = ast.parse(path.read_text())
    for node in tree.body:
        target = node.target if isinstance(node, ast.AnnAssign) else (
            node.targets[0] if isinstance(node, ast.Assign) else None)
        if isinstance(target, ast.Name) and target.id == "DIRBE_25um_BANDPASS":
            wl, weights = node.value.elts
            if isinstance(wl, ast.BinOp):  # [...] * units.micron
                wl = wl.left
            return ast.literal_eval(wl), ast.literal_eval(weights)
    msg = f"DIRBE_25um_BANDPASS not found in {path}"
    raise ValueError(msg)


def main() -> None:
    path = Path(sys.argv[1]) if len(sys.argv) > 1 else REFERENCE_TABLE
    wl, w = bandpass(path)
    assert len(wl) == len(w) == 82
    OUT.write_text(json.dumps({
        "source": "tests/dirbe_tabulated.py:158-328 of zodipy 1.1.3 (DIRBE_25um_BANDPASS)",
        "wavelength_micron": wl, "weight": w}, indent=1) + "\n")
    print(f"wrote {OUT}: {len(wl)} wavelengths {wl[0]}-{wl[-1]} um")


if __name__ == "__main__":
    main()
