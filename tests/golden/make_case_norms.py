"""Extract the known-answer norms of the reference's golden outputs.

Reads data/CASE* of the QWS reference (the printed norms of its test driver,
main.cc:584-795 and :286-312) and writes tests/golden/case_norms.json: for each
case its driver argv and the (name, value) lines that check.sh compares
(check.sh:12-13).  The values are data; nothing of the reference's code is kept.

    python tests/golden/make_case_norms.py [/root/reference]
"""
import json
import os
import re
import sys

ARGV = {  # README:68-80 (lx ly lz lt px py pz pt tol tol_s maxiter maxiter_s)
    "CASE_4111_1111": [4, 1, 1, 1, 1, 1, 1, 1, -1, -1, 6, 50],
    "CASE_4111_1111_HP": [4, 1, 1, 1, 1, 1, 1, 1, -1, -1, 6, 50],
    "CASE0": [32, 6, 4, 3, 1, 1, 1, 1, -1, -1, 6, 50],
    "CASE1": [32, 6, 4, 3, 1, 1, 1, 2, -1, -1, 6, 50],
    "CASE2": [32, 3, 4, 4, 1, 1, 1, 2, -1, -1, 6, 50],
    "CASE3": [32, 3, 6, 4, 1, 1, 1, 2, -1, -1, 6, 50],
    "CASE4": [32, 6, 6, 2, 1, 1, 1, 2, -1, -1, 6, 50],
    "CASE5": [32, 6, 6, 4, 1, 1, 1, 2, -1, -1, 6, 50],
    "CASE6": [32, 6, 6, 8, 1, 1, 1, 2, -1, -1, 6, 50],
    "CASE7": [32, 6, 4, 3, 1, 2, 2, 2, -1, -1, 6, 50],
    "CASE8": [32, 6, 4, 3, 4, 4, 1, 1, -1, -1, 6, 50],
    "CASE9": [32, 6, 4, 3, 1, 1, 4, 4, -1, -1, 6, 50],
}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    pat = re.compile(r" : +[0-9]|2 += +[0-9]")
    out = {}
    for case, argv in ARGV.items():
        lines = []
        section = ""
        with open(os.path.join(ref, "data", case)) as f:
            for line in f:
                if line.startswith("LDDHMC solver"):
                    section = "HP" if "Half Precision" in line else "DD"
                elif line.startswith("test bicgstab"):
                    section = "EO"
                if "git commit hash" in line or not pat.search(line):
                    continue
                name = line.split()[0]
                value = float(line.split()[-1])
                lines.append([name, value, section])
        out[case] = {"argv": argv, "norms": lines}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "case_norms.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
