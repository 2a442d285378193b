"""Writes km_table.json: the reference's 900-entry degree-per-kilometre
longitude table (degree_per_kilometer_longitude_arr, src/include/util.h:73-74
of the reference), as data -- index k covers latitudes [k/10, (k+1)/10).

Run in a container that holds the reference (it reads the header as text and
keeps only the numbers):  python tests/golden/make_km_table.py
"""
import json
import os
import re
import sys

REF = "/root/reference/src/include/util.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "km_table.json")


def main() -> int:
    text = open(REF).read()
    m = re.search(r"degree_per_kilometer_longitude_arr\[\]\s*=\s*\{([^}]*)\}", text)
    if not m:
        print("table not found", file=sys.stderr)
        return 1
    vals = [float(v) for v in re.findall(r"[0-9]+\.[0-9]+", m.group(1))]
    if len(vals) != 900:
        print(f"expected 900 entries, found {len(vals)}", file=sys.stderr)
        return 1
    with open(OUT, "w") as f:
        json.dump({"source": "degree_per_kilometer_longitude_arr, src/include/util.h:73-74",
                   "index": "k = (int)(|latitude| * 10)", "values": vals}, f)
        f.write("\n")
    print(f"wrote {OUT}: {len(vals)} entries")
    return 0


if __name__ == "__main__":
    sys.exit(main())
