"""Writes tests/golden/speed_snps.json: the reference's published speed-test
results (number of segregating mutations after 1,000 generations, compact
interval, run time) per sequence length xL, for the GTX 780M and GTX 980,
read as data from documentation/780M_980_Intel_speed_results.xlsx
(sheet cells A25:D65 and A75:D115).  Run with /root/reference mounted.
"""
import json
import os
import re
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
XLSX = "/root/reference/documentation/780M_980_Intel_speed_results.xlsx"


def cells():
    z = zipfile.ZipFile(XLSX)
    strs = re.findall(r"<t[^>]*>([^<]*)</t>", z.read("xl/sharedStrings.xml").decode())
    sheet = z.read("xl/worksheets/sheet1.xml").decode()
    rows = {}
    for col, row, attr, v in re.findall(r'<c r="([A-Z]+)(\d+)"([^>]*)>(?:<f>[^<]*</f>)?(?:<v>([^<]*)</v>)?', sheet):
        if 't="s"' in attr and v:
            v = strs[int(v)]
        rows.setdefault(int(row), {})[col] = v
    return rows


def block(rows, first, last):
    out = []
    for r in range(first, last + 1):
        c = rows[r]
        out.append(dict(xL=float(c["A"]), compact_interval=int(c["B"]), snps=int(c["C"]), ms=float(c["D"])))
    return out


def main():
    rows = cells()
    data = {
        "source": "documentation/780M_980_Intel_speed_results.xlsx (sheet 780M_980_Intel_results.txt)",
        "scenario": "N=2e5 haploid (F=1), mu=1e-9, L=xL*2e7, neutral, 1000 generations, equilibrium start",
        "780M": block(rows, 25, 65),
        "980": block(rows, 75, 115),
    }
    with open(os.path.join(HERE, "speed_snps.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
