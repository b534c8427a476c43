"""Regenerate the golden fixtures under tests/golden/ from the reference tree.

Reads the block-level checksum tables that the reference's C-port test suite
holds (tests/c/checksums.rs:268-756, themselves ported from upstream
zfp/tests/constants/checksums/*.h) and writes them as data to
checksums.json: {"<dims>d_<type>": [[key1, key2, checksum], ...]}.

Run from the repo root with the reference mounted:
    python tests/golden/make_fixtures.py /root/reference
Only this script reads the reference; tests read the JSON it wrote.
"""
import json
import pathlib
import re
import sys


def main(ref_root):
    src = pathlib.Path(ref_root, "tests", "c", "checksums.rs").read_text()
    tables = {}
    for m in re.finditer(r"static CHECKSUMS_(\d)D_(INT32|INT64|FLOAT|DOUBLE): &\[\(u64, u64, u64\)\] = &\[(.*?)\];",
                         src, re.S):
        dims, ty, body = m.group(1), m.group(2).lower(), m.group(3)
        rows = []
        for t in re.finditer(r"\((0x[0-9a-fA-F_]+),\s*(0x[0-9a-fA-F_]+),\s*(0x[0-9a-fA-F_]+)\)", body):
            rows.append([int(x.replace("_", ""), 16) for x in t.groups()])
        tables[f"{dims}d_{ty}"] = rows
    out = pathlib.Path(__file__).with_name("checksums.json")
    out.write_text(json.dumps(tables, indent=1, sort_keys=True) + "\n")
    print(f"wrote {out} with {sum(len(v) for v in tables.values())} rows in {len(tables)} tables")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
