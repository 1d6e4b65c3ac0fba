"""Regenerate ``kirkland_params.txt`` from the reference's Kirkland fit data.

The reference keeps Kirkland's 12-parameter electron scattering fits
(Kirkland 2010, Appendix C) in ``emsim/assets/atom_params.txt`` as four
lines per element, parameters interleaved ``a1 b1 a2 b2 / a3 b3 c1 d1 /
c2 d2 c3 d3`` (parsed by ``emsim/elem.py:312-354``).  This script rewrites
the same numbers, as the same decimal strings, into one row per element
ordered ``Z chisq a1 a2 a3 b1 b2 b3 c1 c2 c3 d1 d2 d3`` so that
``emsim_amd.elem`` can load them with a single ``numpy.loadtxt``.

Usage: python emsim_amd/data/make_kirkland_table.py [path/to/atom_params.txt]
"""
import sys
from pathlib import Path

DEFAULT_SRC = Path("/root/reference/emsim/assets/atom_params.txt")
DST = Path(__file__).with_name("kirkland_params.txt")


def convert(src: Path) -> list:
    lines = src.read_text().splitlines()
    if len(lines) < 4 * 103:
        raise ValueError(f"{src}: expected 412 lines, found {len(lines)}")
    rows = []
    for z in range(1, 104):
        head, l2, l3, l4 = lines[4 * (z - 1): 4 * z]
        # header reads "Z=  1, chisq= 0.170190"
        z_text, chisq_text = head.split(",")
        if int(z_text.split("=")[1]) != z:
            raise ValueError(f"{src}: element block {z} is out of order")
        chisq = chisq_text.split("=")[1].strip()
        vals = (l2 + " " + l3 + " " + l4).split()
        if len(vals) != 12:
            raise ValueError(f"{src}: element {z} has {len(vals)} parameters")
        a = vals[0:5:2]
        b = vals[1:6:2]
        c = vals[6:11:2]
        d = vals[7:12:2]
        rows.append(" ".join([str(z), chisq] + a + b + c + d))
    return rows


def main(argv):
    src = Path(argv[1]) if len(argv) > 1 else DEFAULT_SRC
    rows = convert(src)
    header = ("# Kirkland (2010) Appendix C electron scattering-factor fits, Z=1..103\n"
              "# columns: Z chisq a1 a2 a3 b1 b2 b3 c1 c2 c3 d1 d2 d3\n"
              "# units: a [1/A], b [1/A^2], c [A], d [A^2]\n")
    DST.write_text(header + "\n".join(rows) + "\n")
    print(f"wrote {DST} ({len(rows)} elements)")


if __name__ == "__main__":
    main(sys.argv)
