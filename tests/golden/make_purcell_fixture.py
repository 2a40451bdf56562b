"""Writes tests/golden/purcell_rows.json from the reference's own Purcell
spectra fE.dat / fdip.dat (5001 rows of omega, amplitude, phase each; they
are data files, read with numpy.loadtxt, nothing is executed).  Only the
rows of the reference's plot range 0.4 <= 2pi/omega <= 0.7 (purcell.gp:2)
are kept, with the value purcell.gp's expression gives for each row.

    python tests/golden/make_purcell_fixture.py /root/reference
"""
import json
import os
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
E = np.loadtxt(os.path.join(ref, "fE.dat"), comments="#")
D = np.loadtxt(os.path.join(ref, "fdip.dat"), comments="#")
assert E.shape == D.shape == (5001, 3)
assert np.allclose(E[:, 0], D[:, 0])
w = E[:, 0]
with np.errstate(divide="ignore"):
    lam = np.where(w > 0, 2 * np.pi / w, np.inf)
rows = np.where((lam >= 0.4) & (lam <= 0.7))[0]
out = []
for r in rows:
    om, AE, pE = E[r]
    _, Ad, pd = D[r]
    P = (AE * np.cos(pE) * Ad * np.sin(pd) - AE * np.sin(pE) * Ad * np.cos(pd)) * om / 2 / (Ad ** 2 * om ** 4 / 3) + 1
    out.append(dict(row=int(r), omega=om, AE=AE, phE=pE, Ad=Ad, phd=pd, P=P))
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "purcell_rows.json"), "w") as f:
    json.dump(dict(source="fE.dat, fdip.dat (reference data files), rows with 0.4<=2pi/omega<=0.7",
                   formula="purcell.gp:2", rows=out), f, indent=1)
print(len(out), "rows")
