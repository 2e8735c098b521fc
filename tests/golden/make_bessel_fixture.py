"""Writes bessel_mpmath.json: J0, Y0, J1, Y1 at 30 significant digits (mpmath)
on a node set spanning every branch of the J/Y evaluation (x <= 0 excluded):
limiting forms (< 1e-8), ascending series (< 8), asymptotic P/Q forms
(8 .. 1e6) and the leading asymptotic term (> 1e6).

Run: python make_bessel_fixture.py  (writes next to this script)
"""
import json
import os

import mpmath as mp

XS = [1e-12, 1e-9, 9.999e-9, 1.0001e-8, 1e-6, 1e-3, 0.1, 0.5, 1.0, 2.0, 2.404825557695773, 3.0, 3.14159265358979,
      3.831705970207512, 5.0, 7.0, 7.999, 7.9999999, 8.0, 8.0000001, 8.001, 10.0, 15.0, 25.0, 40.0, 63.5, 100.0,
      250.0, 540.0, 1000.0, 12345.678, 999999.0, 1000001.0, 1e7]


def main():
    mp.mp.dps = 40
    rows = []
    for x in XS:
        xm = mp.mpf(x)
        rows.append({"x": x,
                     "J0": mp.nstr(mp.besselj(0, xm), 30), "Y0": mp.nstr(mp.bessely(0, xm), 30),
                     "J1": mp.nstr(mp.besselj(1, xm), 30), "Y1": mp.nstr(mp.bessely(1, xm), 30)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bessel_mpmath.json")
    with open(path, "w") as f:
        json.dump({"source": "mpmath %s, 40-digit working precision" % mp.__version__, "rows": rows}, f, indent=1)


if __name__ == "__main__":
    main()
