"""Write tests/golden/golden.json: mpmath reference values (30 digits) that pin
the oracle and the kernels' special functions and Green's functions.

The Green's functions are evaluated here in textbook forms written
independently of the reference's expressions:
  3D dyadic  e^{ikR}/(4 pi R) [(1 + i/kR - 1/(kR)^2) I + (-1 - 3i/kR + 3/(kR)^2) R̂R̂]
  3D scalar  e^{ikR}/(4 pi R)
  2D TM      (i/4) H0(kR)
  2D TE      (i/8) [(H0(kR) - H2(kR)) I + 2 H2(kR) R̂R̂]   (H1(x)/x = (H0 + H2)/2)
and the incident fields from their closed forms.
Run: python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 30
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")


def c(z):
    z = mp.mpc(z)
    return [float(z.real), float(z.imag)]


def bessel_cases():
    xs = [1e-8, 1e-5, 1e-3, 0.01, 0.1, 0.5, 0.8935769662791675, 1.0, 2.0, 2.404825557695773, 3.0,
          3.8317059702075125, 4.5, 5.520078110286311, 6.0, 7.0, 7.999999, 8.0, 8.000001, 9.0, 10.0, 12.5,
          15.0, 20.0, 30.0, 50.0, 75.0, 100.0, 150.0, 222.0, 500.0, 1000.0, 1e4]
    return [{"x": x, "j0": float(mp.besselj(0, x)), "y0": float(mp.bessely(0, x)),
             "j1": float(mp.besselj(1, x)), "y1": float(mp.bessely(1, x))} for x in xs]


def green3d(r, k):
    R = mp.sqrt(sum(mp.mpf(v) ** 2 for v in r))
    kr = k * R
    a = 1 + 1j / kr - 1 / kr ** 2
    b = -1 - 3j / kr + 3 / kr ** 2
    ph = mp.exp(1j * kr) / (4 * mp.pi * R)
    return [[c(ph * ((a if p == q else 0) + b * mp.mpf(r[p]) * mp.mpf(r[q]) / R ** 2)) for q in range(3)]
            for p in range(3)]


def green3ds(r, k):
    R = mp.sqrt(sum(mp.mpf(v) ** 2 for v in r))
    return c(mp.exp(1j * k * R) / (4 * mp.pi * R))


def green2d_tm(r, k):
    R = mp.sqrt(sum(mp.mpf(v) ** 2 for v in r))
    return c(0.25j * mp.hankel1(0, k * R))


def green2d_te(r, k):
    R = mp.sqrt(sum(mp.mpf(v) ** 2 for v in r))
    h0, h2 = mp.hankel1(0, k * R), mp.hankel1(2, k * R)
    return [[c(mp.mpf(1) / 8 * 1j * ((h0 - h2) * (1 if p == q else 0)
                                       + 2 * h2 * mp.mpf(r[p]) * mp.mpf(r[q]) / R ** 2)) for q in range(2)]
            for p in range(2)]


def green_cases():
    r3 = [[0.3, -0.2, 0.1], [1e-3, 2e-3, -1.5e-3], [1.2, 0.4, -0.9], [0.0, 0.0, 0.05], [-2.0, 1.0, 0.5]]
    r2 = [[0.3, -0.2], [1e-3, 2e-3], [1.2, 0.4], [0.0, 0.05], [-2.0, 1.0], [5.0, -3.0]]
    ks = [1.0, 8.46, 17.59, 35.0]
    out = {"3d_vector": [], "3d_scalar": [], "2d_tm": [], "2d_te": []}
    for k in ks:
        for r in r3:
            out["3d_vector"].append({"r": r, "k0": k, "G": green3d(r, k)})
            out["3d_scalar"].append({"r": r, "k0": k, "G": green3ds(r, k)})
        for r in r2:
            out["2d_tm"].append({"r": r, "k0": k, "G": green2d_tm(r, k)})
            out["2d_te"].append({"r": r, "k0": k, "G": green2d_te(r, k)})
    return out


def beam(rpar, rperp, k, w):
    a = 2 * rperp / (w * w * k)
    return mp.exp(1j * rpar * k - rperp ** 2 / (w * w * (1 + 1j * a))) / mp.sqrt(1 + 1j * a)


def source_cases():
    out = {"2d": [], "3d": []}
    k, w = 17.3, 0.2
    for (x, y) in [(0.1, 0.2), (-0.3, 0.05), (0.0, 0.0), (0.25, -0.4)]:
        for th in [0.0, 0.7, 2.5, 4.0]:
            rpar = mp.mpf(x) * mp.cos(th) + mp.mpf(y) * mp.sin(th)
            rperp = -mp.mpf(x) * mp.sin(th) + mp.mpf(y) * mp.cos(th)
            out["2d"].append({"x": [x, y], "theta": th, "k0": k, "w": w, "plane": c(mp.exp(1j * k * rpar)),
                              "beam": c(beam(rpar, rperp, k, w))})
    for (x, y, z) in [(0.1, 0.2, -0.1), (-0.3, 0.05, 0.2), (0.0, 0.0, 0.0)]:
        for u in [(0.0, 0.0, 1.0), (0.6, 0.0, 0.8), (-0.48, 0.6, 0.64)]:
            rpar = mp.mpf(x) * u[0] + mp.mpf(y) * u[1] + mp.mpf(z) * u[2]
            e = [mp.mpf(x) - rpar * u[0], mp.mpf(y) - rpar * u[1], mp.mpf(z) - rpar * u[2]]
            rperp = mp.sqrt(sum(v * v for v in e))
            out["3d"].append({"x": [x, y, z], "u": list(u), "k0": k, "w": w, "plane": c(mp.exp(1j * k * rpar)),
                              "beam": c(beam(rpar, rperp, k, w))})
    return out


def main():
    data = {"bessel": bessel_cases(), "green": green_cases(), "sources": source_cases()}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
