#!/usr/bin/env python3
"""Extract the screening-count fixture of the reference's tests/test_screening.py.

That test (test_w2_631gd, tests/test_screening.py:234-382 of gms_libERI)
sends a water-dimer 6-31G* shell layout and GAMESS Schwarz integrals over MDI
and compares per-class surviving shell-quartet counts with a target array
(:380, alternative target commented at :381).  Its harness needs the MDI
library (and asks for an MDI command the plugin never registered), so only its
data is used: the literal arrays are read from the source text here and
written to ref_screening_w2_631gd.json.  Shells 17-20 of the test have
ang_mom 0 (padding) and are dropped; Schwarz entries of the 16 real shells are
the first 16*17/2 packed entries.
"""
from __future__ import annotations

import ast
import json
import os
import re
from pathlib import Path

REF = Path(os.environ.get("LIBERI_REFERENCE", "/root/reference"))
OUT = Path(__file__).resolve().parent / "ref_screening_w2_631gd.json"

# class labels of the target array (comment at tests/test_screening.py:379)
# mapped to the class names used by the library (dispatch names of src/liberi_driver.F90)
LABELS = {"0001": "int0001", "0011": "int0011", "0111": "int0111", "0002": "int0002", "0012": "int0012",
          "0112": "int0112", "1112": "int2111", "0022": "int0022", "0122": "int0122", "1122": "int2211",
          "0222": "int2220", "1222": "int2221", "0102": "int0102", "0212": "int2120"}


def grab(text: str, name: str, start: int):
    m = re.compile(rf"^\s*{name}\s*=\s*(\[.*?\])", re.S | re.M).search(text, start)
    return ast.literal_eval(m.group(1).replace("\n", " ")), text[: m.start()].count("\n") + 1


def main():
    src = REF / "tests" / "test_screening.py"
    text = src.read_text()
    start = text.index("def test_w2_631gd")
    ang, l1 = grab(text, "ang_mom", start)
    sw, l2 = grab(text, "schwrz_int", start)
    m = re.compile(r"^\s*target_array = np\.array\((\[.*?\])\)", re.M).search(text, start)
    target = ast.literal_eval(m.group(1))
    tline = text[: m.start()].count("\n") + 1
    alt = re.compile(r"^\s*#\s*target_array = np\.array\((\[.*?\])\)", re.M).findall(text[m.end():])[0]
    alt = ast.literal_eval(alt)
    order = re.search(r"#(0001,[0-9,]+)", text[start:]).group(1).split(",")
    nreal = sum(1 for a in ang if a > 0)
    data = {
        "source": f"gms_libERI tests/test_screening.py:{l1} (ang_mom), :{l2} (schwrz_int), :{tline} (target), "
                  f":{tline + 1} (alternative target)",
        "ang_mom": ang[:nreal],
        "schwrz": sw[: nreal * (nreal + 1) // 2],
        "classes": [LABELS[o] for o in order],
        "target": target,
        "alternative_target": alt,
    }
    OUT.write_text(json.dumps(data, indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
