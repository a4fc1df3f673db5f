"""Writes besselk_golden.json: K_nu(x) at 40 significant digits (mpmath) on a
grid covering the engine's domain (nu in [0.01, 3], x in [1e-8, 700]).
Run: python tests/golden/make_besselk_golden.py"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
NUS = [0.01, 0.1, 0.25, 0.3, 0.5, 0.7, 0.8, 0.9999, 1.0, 1.25, 1.5, 1.7, 2.0, 2.25, 2.5, 2.75,
       2.9, 3.0]
XS = [1e-8, 1e-5, 1e-3, 0.01, 0.05, 0.1, 0.3, 0.5, 0.9, 1.0, 1.5, 1.9, 1.99, 2.0, 2.01, 2.5,
      3.0, 4.0, 6.0, 10.0, 20.0, 50.0, 100.0, 300.0, 600.0, 700.0]
rows = []
for nu in NUS:
    for x in XS:
        rows.append([nu, x, mp.nstr(mp.besselk(nu, x), 25)])
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "besselk_golden.json")
with open(out, "w") as f:
    json.dump({"source": "mpmath %s besselk, dps=40" % mp.__version__, "rows": rows}, f, indent=0)
print(out, len(rows))
