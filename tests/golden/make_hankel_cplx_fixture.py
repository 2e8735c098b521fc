"""Writes hankel_cplx_mpmath.json: H0^(1)(z), H1^(1)(z) at 50 digits (mpmath)
through H0(z) = -(2i/pi) K0(-iz), H1(z) = -(2/pi) K1(-iz) (A&S 9.6.4; mpmath's
besselk, which stays accurate where J + iY cancels for large Im z), checked
against mpmath's hankel1 where that is accurate, for complex z in the closed
upper half-plane -- both sides of |z| = 2 (the
series / continued-fraction switch of K0(-iz), K1(-iz)), near the real axis,
the negative real axis, large imaginary parts and large |z|.

Run:  python make_hankel_cplx_fixture.py > hankel_cplx_mpmath.json
"""
import json

import mpmath as mp


def main():
    mp.mp.dps = 50
    pts = []
    for r in (1e-6, 0.05, 0.7, 1.5, 1.99, 2.01, 2.6, 4.0, 7.5, 13.0, 30.0, 120.0, 900.0):
        for th in (0.0, 1e-7, 0.3, 0.9, 1.5707, 2.2, 3.0, 3.14159265358979):
            z = mp.mpc(r * mp.cos(th), r * mp.sin(th))
            pts.append(z)
    rows = []
    for z in pts:
        w = -1j * z
        h0 = -2j / mp.pi * mp.besselk(0, w)
        h1 = -2 / mp.pi * mp.besselk(1, w)
        if abs(z.imag) < 20:
            for h, ref in ((h0, mp.hankel1(0, z)), (h1, mp.hankel1(1, z))):
                assert abs(h - ref) <= mp.mpf(10) ** -30 * abs(ref), (z, h, ref)
        rows.append({"re": float(z.real), "im": float(z.imag),
                     "h0": [mp.nstr(h0.real, 25), mp.nstr(h0.imag, 25)],
                     "h1": [mp.nstr(h1.real, 25), mp.nstr(h1.imag, 25)]})
    print(json.dumps({"source": "mpmath hankel1, 50 digits", "rows": rows}, indent=0))


if __name__ == "__main__":
    main()
