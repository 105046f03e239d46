"""Write the golden fixtures under tests/golden/ from the curtree sources.

Run in a container where the reference tree is mounted (default
/root/reference); the fixtures it writes are plain data and are committed, so
the tests never read the reference at run time.

Fixtures:
  dilate_table.json   the 256-entry dilation table and the 256-entry inverse
                      table of src/z_order.cuh:20-42 and :44-60 (Raman & Wise),
                      parsed from the source text.  These are the reference's
                      own known-answer values for the Morton codec.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_table(text, name):
    m = re.search(name + r"\s*\[\s*256\s*\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit(f"table {name} not found")
    vals = [int(v, 16) for v in re.findall(r"0x[0-9A-Fa-f]+", m.group(1))]
    if len(vals) != 256:
        raise SystemExit(f"table {name}: {len(vals)} entries")
    return vals


def main(ref_root="/root/reference"):
    src = os.path.join(ref_root, "src", "z_order.cuh")
    with open(src) as f:
        text = f.read()
    out = {
        "source": "src/z_order.cuh:20-42 (dilate_table), :44-60 (undilate_table)",
        "dilate_table": parse_table(text, "dilate_table"),
        "undilate_table": parse_table(text, "undilate_table"),
    }
    with open(os.path.join(HERE, "dilate_table.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote dilate_table.json")


if __name__ == "__main__":
    main(*sys.argv[1:])
