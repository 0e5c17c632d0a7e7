#!/usr/bin/env python3
"""Listing cycle lines (ecrit.F FORMAT 1100) taken from the Engine listings (*_0001.out) that the
reference's QA tests keep as their expected output: up to 12 lines per file, evenly spaced over
its cycle lines, plus every line with a three-digit exponent or a negative value.  The lines are
data: the test re-prints each from its parsed fields and compares the text.

Run in the build container:  python tests/golden/make_listing_lines.py
"""
import glob
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def main():
    out = []
    for f in sorted(glob.glob(os.path.join(REF, "**", "*_0001.out"), recursive=True)):
        lines = []
        for line in open(f, errors="replace"):
            fl = line.split()
            if len(fl) == 13 and fl[5].endswith("%"):
                lines.append(line.rstrip("\n"))
        if not lines:
            continue
        n = len(lines)
        pick = sorted({round(i * (n - 1) / 11) for i in range(12)} if n > 12 else range(n))
        keep = [lines[i] for i in pick]
        keep += [s for s in lines if (("-1" in s.split("%")[1] and "E" not in s.split()[6]) or " -" in s[40:])
                 and s not in keep][:8]
        out.append("# " + os.path.relpath(f, REF))
        out.extend(keep)
    path = os.path.join(HERE, "listing_lines.txt")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    print(path, sum(1 for s in out if not s.startswith("#")), "lines")


if __name__ == "__main__":
    main()
