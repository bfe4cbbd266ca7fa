"""Write tests/golden/dirbe_kat.json: the DIRBE known-answer test inputs and outputs.

* The table of the reference's test (tests/dirbe_tabulated.py:3-156: the DIRBE
  Zodiacal Light Prediction Software's emission for 10 days of 1990 and 10
  ecliptic directions in each of the 10 DIRBE bands) is read as text with
  ``ast`` -- nothing of the reference is imported or executed.
* Earth's heliocentric ecliptic position on each day comes from the published
  analytic ephemeris "Keplerian Elements for Approximate Positions of the Major
  Planets" (E. M. Standish, JPL Solar System Dynamics, Table 1, valid
  1800-2050; Earth-Moon barycentre, mean ecliptic and equinox of J2000). Its
  error (< 1e-4 AU here, plus 3e-5 AU from barycentre to Earth) moves the
  emission by < 1e-3 relative, under the test's 1 % tolerance. The reference
  takes the position from astropy's "builtin" ephemeris
  (zodipy/bodies.py:53-62); astropy is not importable here.

    python tests/golden/make_dirbe_kat.py [/root/reference/tests/dirbe_tabulated.py]
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

import numpy as np

REFERENCE_TABLE = Path("/root/reference/tests/dirbe_tabulated.py")
JD_1990_01_01 = 2447892.5  # 1990-01-01T00:00 (UTC; the 57 s to TT do not matter here)

# Standish, Table 1 (1800 AD - 2050 AD), EM Bary: value at J2000 and rate per century.
EMB = {
    "a": (1.00000261, 0.00000562),
    "e": (0.01671123, -0.00004392),
    "I": (-0.00001531, -0.01294668),
    "L": (100.46457166, 35999.37244981),
    "varpi": (102.93768193, 0.32327364),
    "Omega": (0.0, 0.0),
}


def emb_heliocentric(jd: float) -> np.ndarray:
    """Heliocentric J2000 ecliptic position [AU] by the recipe of the same document."""
    T = (jd - 2451545.0) / 36525.0
    el = {k: v0 + v1 * T for k, (v0, v1) in EMB.items()}
    a, e = el["a"], el["e"]
    inc, L, varpi, Omega = (np.radians(el[k]) for k in ("I", "L", "varpi", "Omega"))
    omega = varpi - Omega
    M = (L - varpi + np.pi) % (2 * np.pi) - np.pi
    E = M + e * np.sin(M)
    for _ in range(50):  # Newton on Kepler's equation E - e sin E = M
        dE = (M - (E - e * np.sin(E))) / (1 - e * np.cos(E))
        E += dE
        if abs(dE) < 1e-15:
            break
    xp = a * (np.cos(E) - e)
    yp = a * np.sqrt(1 - e * e) * np.sin(E)
    co, so, cO, sO = np.cos(omega), np.sin(omega), np.cos(Omega), np.sin(Omega)
    ci, si = np.cos(inc), np.sin(inc)
    x = (co * cO - so * sO * ci) * xp + (-so * cO - co * sO * ci) * yp
    y = (co * sO + so * cO * ci) * xp + (-so * sO + co * cO * ci) * yp
    z = (so * si) * xp + (co * si) * yp
    return np.array([x, y, z])


def read_table(path: Path) -> dict:
    """Literal assignments DAYS, LON, LAT, TABULATED_DIRBE_EMISSION of the reference file."""
    tree = ast.parse(path.read_text())
    want = {"DAYS", "LON", "LAT", "TABULATED_DIRBE_EMISSION"}
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            name = getattr(node.targets[0], "id", None)
            if name in want:
                out[name] = ast.literal_eval(node.value)
    missing = want - set(out)
    if missing:
        msg = f"{path}: missing {sorted(missing)}"
        raise SystemExit(msg)
    return out


def main() -> None:
    src = Path(sys.argv[1]) if len(sys.argv) > 1 else REFERENCE_TABLE
    tab = read_table(src)
    earth = [emb_heliocentric(JD_1990_01_01 + d - 1).tolist() for d in tab["DAYS"]]
    out = {
        "source": "tests/dirbe_tabulated.py of zodipy (DIRBE Zodiacal Light Prediction Software)",
        "earth_ephemeris": "JPL approximate Keplerian elements (Standish), EM barycentre",
        "days_since_1990_01_00": tab["DAYS"],
        "lon_deg": tab["LON"],
        "lat_deg": tab["LAT"],
        "earth_au": earth,
        "emission_mjy_sr": {str(k): v for k, v in tab["TABULATED_DIRBE_EMISSION"].items()},
    }
    path = Path(__file__).with_name("dirbe_kat.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
