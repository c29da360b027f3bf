"""Write the committed golden fixtures from the reference's own data files.

Runs only in the build container, where the reference checkout exists
(default /root/reference, or $HWTREE_REFERENCE).  It copies data, never code:
each fixture is an input file in the reference's text format
(common/Options.hpp:87-94, common/Yield.hpp:34-35, common/Arrays.hpp:15-46)
and the matching golden outputs written by the reference's seq binary:
FP64 (seq/compute.sh:8,18,33 -> data/out/, fixture `.out`) and FP32 (the
default `real = float` build -> data/out32/, fixture `.out32`).  Tokens are
copied verbatim so the numbers parse to exactly the values the reference read
or wrote.

Slices: the small files are taken whole; each 65,536-option set is sliced to
every 64th option (1,024 rows), the 60,000-option set to every 60th.
"""
import os
import sys
from pathlib import Path

REF = Path(os.environ.get("HWTREE_REFERENCE", "/root/reference"))
OUT = Path(__file__).resolve().parent

WHOLE = ["book", "options-1000", "0_UNIFORM_1000", "1_RAND_1000",
         "2_RANDCONSTHEIGHT_1000", "3_RANDCONSTWIDTH_1000"]
SLICED = {"4_SKEWED": 64, "5_SKEWEDCONSTHEIGHT": 64, "6_SKEWEDCONSTWIDTH": 64,
          "options-60000": 60}


def read_token_lines(path):
    lines = []
    for raw in path.read_text().splitlines():
        raw = raw.strip()
        if not raw:
            continue
        assert raw[0] == "[" and raw[-1] == "]", path
        lines.append([t.strip() for t in raw[1:-1].split(",")])
    return lines


def write_token_lines(path, lines):
    path.write_text("".join("[" + ", ".join(toks) + "]\n" for toks in lines))


def main():
    if not (REF / "data").is_dir():
        sys.exit(f"reference data not found under {REF}")
    (OUT / "yield.in").write_text((REF / "data" / "yield.in").read_text())
    # the data files are the reference's, under its ISC licence: its notice
    # travels with them
    (OUT / "LICENSE-reference-data.txt").write_text(
        "The .in, .out and .out32 files in this directory and width_table.json are\n"
        "taken from MartinMetaksov/diku.OptionsPricing (data/, data/out/, data/out32/,\n"
        "data-generator/RandomDataGenerator.cpp) under the following licence.\n\n"
        + (REF / "LICENSE.txt").read_text())
    for name in WHOLE:
        write_token_lines(OUT / f"{name}.in", read_token_lines(REF / "data" / f"{name}.in"))
        write_token_lines(OUT / f"{name}.out", read_token_lines(REF / "data" / "out" / f"{name}.out"))
        write_token_lines(OUT / f"{name}.out32", read_token_lines(REF / "data" / "out32" / f"{name}.out"))
    for name, step in SLICED.items():
        src_in = read_token_lines(REF / "data" / f"{name}.in")
        src_out = read_token_lines(REF / "data" / "out" / f"{name}.out")
        src_out32 = read_token_lines(REF / "data" / "out32" / f"{name}.out")
        n = len(src_out[0])
        assert all(len(t) == n for t in src_in) and len(src_out32[0]) == n, name
        keep = list(range(0, n, step))
        write_token_lines(OUT / f"{name}_s{step}.in", [[t[i] for i in keep] for t in src_in])
        write_token_lines(OUT / f"{name}_s{step}.out", [[src_out[0][i] for i in keep]])
        write_token_lines(OUT / f"{name}_s{step}.out32", [[src_out32[0][i] for i in keep]])
    # Known-answer table for the tree width: the data generator's pairs of
    # reversion rate -> width at 12 steps/yr (data-generator/RandomDataGenerator.cpp:54-55).
    import json
    import re
    src = (REF / "data-generator" / "RandomDataGenerator.cpp").read_text()
    rates = [float(x) for x in re.search(r"reversionRates = \{([^}]*)\}", src).group(1).split(",")]
    widths = [int(x) for x in re.search(r"widths = \{([^}]*)\}", src).group(1).split(",")]
    (OUT / "width_table.json").write_text(json.dumps({"term_steps": 12, "term_unit": 365,
                                                      "a": rates, "width": widths}))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
