"""Write the replay fixtures from the reference's OPT_CYCLE logs.

The reference ships two logs of method-'G' cyclic optimisation runs on the
HD+ 1500-function bases (example_output/out_1_1500, out_2_1500: roots 1 and
2).  Each complete record of a log is one optimised function:

      Optimizing function   809
     Nonlinear parameters before optimization:
       809:  <6 parameters, D23.16>
     Number of energy/gradient evaluations          31           3
     E=-0.5978979685881810D+00
     Nonlinear parameters after optimization:
       809:  <6 parameters>

(the E printed is the energy with the function's "after" parameters and
every earlier record applied).  The logs stop inside a record; the
reference's example_input/inout_{1,2}_1500.txt are the states after the
last complete record.  The log's header also prints the CURRENT_ENERGY of
the state before the first record.

Output: replay_{1,2}_1500.npz with func (1-based function index), before,
after (records x 6), E (records), nevals (records x 2: the evaluation
counts the log prints), E_initial (the header's CURRENT_ENERGY).  Run from
the repository root: python tests/golden/replay/make_replay.py
"""
from __future__ import annotations

import os
import re

import numpy as np

REF = "/root/reference/example_output"
HERE = os.path.dirname(os.path.abspath(__file__))


def fd(s):
    return float(s.replace("D", "E"))


def params_line(line):
    # "   809:  0.1521314192029310D+01-0.1455248340858359D+01 ..." (D23.16 fields)
    body = line.split(":", 1)[1]
    vals = re.findall(r"[-+ ]?\d\.\d{16}D[-+]\d{2}", body)
    return [fd(v) for v in vals]


def parse(path):
    lines = open(path).read().splitlines()
    e0 = None
    for ln in lines:
        if ln.strip().startswith("CURRENT_ENERGY"):
            e0 = fd(ln.split()[1])
            break
    recs = []
    i = 0
    while i < len(lines):
        m = re.match(r"\s*Optimizing function\s+(\d+)", lines[i])
        if m and i + 5 < len(lines) and lines[i + 4].strip().startswith("E=") and \
                lines[i + 5].strip().startswith("Nonlinear parameters after"):
            f = int(m.group(1))
            before = params_line(lines[i + 2])
            nev = [int(v) for v in lines[i + 3].split()[-2:]]
            E = fd(lines[i + 4].strip()[2:])
            after = params_line(lines[i + 6])
            assert len(before) == 6 and len(after) == 6, (path, i)
            recs.append((f, before, after, E, nev))
            i += 7
        else:
            i += 1
    return e0, recs


def main():
    for tag in ("1", "2"):
        e0, recs = parse(os.path.join(REF, "out_%s_1500" % tag))
        np.savez_compressed(
            os.path.join(HERE, "replay_%s_1500.npz" % tag),
            func=np.array([r[0] for r in recs], np.int32),
            before=np.array([r[1] for r in recs]),
            after=np.array([r[2] for r in recs]),
            E=np.array([r[3] for r in recs]),
            nevals=np.array([r[4] for r in recs], np.int32),
            E_initial=np.array(e0))
        print(tag, len(recs), "records, E_initial", e0, "last E", recs[-1][3])


if __name__ == "__main__":
    main()
