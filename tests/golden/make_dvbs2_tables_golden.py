"""Writes tests/golden/dvbs2_tables.json: for every DVB-S2 code of the reference, the parameter q
and a SHA-256 of its address table, read from the reference's own data (src/DvbS2.cpp:26-1593,
the `parameter_` and `index_` initialisers), so that the library's tables are pinned to the
reference's numbers without a copy of them here.

    python tests/golden/make_dvbs2_tables_golden.py [/root/reference]

Canonical form hashed: the table's rows in order, each row's addresses sorted ascending and
joined by ',', rows joined by ';' (sorting makes the hash independent of the order the
addresses are listed in, which does not change the matrix).  Also recorded: the number of rows
(groups of 360 information bits) and each row's length.
"""
import hashlib
import json
import os
import re
import sys

RATES = {64800: ["1/4", "1/3", "2/5", "1/2", "3/5", "2/3", "3/4", "4/5", "5/6", "8/9", "9/10"],
         16200: ["1/4", "1/3", "2/5", "1/2", "3/5", "2/3", "3/4", "4/5", "5/6", "8/9"]}


def nested(text):
    """Parses a brace initialiser of integers into nested lists."""
    stack, root, num = [], None, ""
    for ch in text:
        if ch.isdigit():
            num += ch
            continue
        if num:
            stack[-1].append(int(num))
            num = ""
        if ch == "{":
            new = []
            if stack:
                stack[-1].append(new)
            else:
                root = new
            stack.append(new)
        elif ch == "}":
            stack.pop()
    return root


def initialiser(src, name):
    m = re.search(r"Ldpc::DvbS2::" + name + r"\s*=\s*", src)
    i = m.end()
    depth = 0
    for j in range(i, len(src)):
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                return src[i:j + 1]
    raise ValueError(name)


def canonical(table):
    return ";".join(",".join(str(a) for a in sorted(row)) for row in table)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = open(os.path.join(ref, "src", "DvbS2.cpp")).read()
    params = nested(initialiser(src, "parameter_"))
    index = nested(initialiser(src, "index_"))
    codes = {}
    for li, n in enumerate((64800, 16200)):
        for ri, rate in enumerate(RATES[n]):
            table = index[li][ri]
            codes[f"{n} {rate}"] = {
                "q": params[li][ri],
                "groups": len(table),
                "row_lengths": [len(r) for r in table],
                "table_sha256": hashlib.sha256(canonical(table).encode()).hexdigest(),
            }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dvbs2_tables.json")
    with open(out, "w") as f:
        json.dump({"source": "src/DvbS2.cpp:26-1593 (parameter_, index_)",
                   "canonical": "rows in table order; addresses sorted ascending, ',' within a row, ';' between rows",
                   "codes": codes}, f, indent=1)
    print("wrote", out, len(codes), "codes")


if __name__ == "__main__":
    main()
