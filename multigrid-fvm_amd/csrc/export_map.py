#!/usr/bin/env python3
"""Write a linker version script that exports exactly the C entry points the
given headers declare and hides every other symbol of the library (its C++
launchers, the set-up helpers, HIP's registration symbols).

The FP64 library and the sensitivity library are built from the same
sources with different scalar types; hiding everything but the declared C
ABI keeps one library's internal calls from binding to the other's code
when both are loaded into one process.

usage: export_map.py OUT HEADER...
"""
import re
import sys


def declared(paths):
    names = set()
    for p in paths:
        txt = open(p).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        names.update(re.findall(r"\b(mgf_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def main():
    out, headers = sys.argv[1], sys.argv[2:]
    names = declared(headers)
    with open(out, "w") as f:
        f.write("{\n  global:\n")
        for n in names:
            f.write(f"    {n};\n")
        f.write("  local:\n    *;\n};\n")


if __name__ == "__main__":
    main()
