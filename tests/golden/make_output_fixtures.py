"""Copies excerpts of the reference's shipped output files, verbatim, into
tests/golden/output_formats/ (run in the development container, where
/root/reference exists). They pin the text layout of the writers in
sph2d_output.hip: header lines, ZONE lines and the F20.10 / F20.8 rows.

Sources (mini-walker/SPH-2D, Taylor_Green_Vortex/, the shipped 100x100 run):
  Final_Doamin_Results.dat  header + the first 200 particle rows (8F20.10 of 6 values)
  Free_Surface.dat          header + the first 4 ZONE lines (all zones are empty)
  Kinetic_energy.dat        first 40 rows (format 310 = 20F20.8)
  Norm_Pressure.dat         first 40 rows
  Wave_front_results.dat    all 500 rows (ifort list-directed REAL(8): E form below 0.1)
  Initial_Post_Proceeding_Grid.dat / _Node.dat   the first 8 lines, the line count and
                            the SHA-256 of the whole file (10201 nodes, 10000 quads)
Only data lines are copied; no reference source is.
"""
import hashlib
import json
import os

REF = os.environ.get("SPH2D_REFERENCE", "/root/reference") + "/Taylor_Green_Vortex"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "output_formats")
EXCERPTS = {"Final_Doamin_Results.dat": 203, "Free_Surface.dat": 6, "Kinetic_energy.dat": 40,
            "Norm_Pressure.dat": 40, "Wave_front_results.dat": 500}


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, nlines in EXCERPTS.items():
        with open(os.path.join(REF, name), "rb") as f:
            lines = f.read().split(b"\n")[:nlines]
        with open(os.path.join(OUT, name + ".head"), "wb") as f:
            f.write(b"\n".join(lines) + b"\n")
        print(name, len(lines), "lines")
    digest = {}
    for name in ("Initial_Post_Proceeding_Grid.dat", "Initial_Post_Proceeding_Node.dat"):
        with open(os.path.join(REF, name), "rb") as f:
            data = f.read()
        digest[name] = {"sha256": hashlib.sha256(data).hexdigest(), "lines": data.count(b"\n"),
                        "head": [ln.decode() for ln in data.split(b"\n")[:8]]}
    with open(os.path.join(OUT, "initial_grid.json"), "w") as f:
        json.dump(digest, f, indent=1)
    print("initial grid digests written")


if __name__ == "__main__":
    main()
