"""Generate bessel_coeffs.h: Chebyshev fits for FP64 J0, Y0, J1, Y1.

Method (Abramowitz & Stegun 9.1.10-9.1.13 and 9.2.5-9.2.6, fitted here):

* 0 <= x <= 8, with t = x^2/64 and s = 2t - 1:
    J0(x) = F_J0(s)
    Y0(x) = (2/pi) ln(x/2) J0(x) + F_R0(s)        (A&S 9.1.13: the remainder is entire in x^2)
    J1(x) = x F_J1(s)
    Y1(x) = (2/pi)(ln(x/2) J1(x) - 1/x) + x F_R1(s)   (A&S 9.1.11)
* x > 8, with z = 8/x, u = z^2 and s = 2u - 1 (Hankel's modulus/phase form, A&S 9.2.5-9.2.6):
    J_n = sqrt(2/(pi x)) (P_n cos chi_n - Q_n sin chi_n)
    Y_n = sqrt(2/(pi x)) (P_n sin chi_n + Q_n cos chi_n),  chi_n = x - (2n+1) pi/4
  with P_n = F_Pn(s) and Q_n = z F_Qn(s).

Each F is interpolated at Chebyshev nodes with mpmath at 50 digits and the
series is cut where the coefficients fall below 1e-19 of the function scale.
Run:  python gen_bessel.py > bessel_coeffs.h
"""
import mpmath as mp

mp.mp.dps = 50
NODES = 60


def cheb_fit(f, n=NODES):
    xs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    fs = [f(x) for x in xs]
    coef = []
    for j in range(n):
        s = mp.fsum(fs[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / n) for k in range(n))
        coef.append(2 * s / n)
    coef[0] /= 2
    return coef


def trim(coef, tol=mp.mpf('1e-19')):
    scale = max(abs(c) for c in coef)
    last = 0
    for i, c in enumerate(coef):
        if abs(c) > tol * scale:
            last = i
    return coef[: last + 1]


def small(fn):
    def f(s):
        t = (s + 1) / 2
        x = 8 * mp.sqrt(t)
        return fn(x)
    return f


def j0(x):
    return mp.besselj(0, x)


def r0(x):
    if x == 0:
        return 2 / mp.pi * mp.euler
    return mp.bessely(0, x) - 2 / mp.pi * mp.log(x / 2) * mp.besselj(0, x)


def j1x(x):
    if x == 0:
        return mp.mpf(1) / 2
    return mp.besselj(1, x) / x


def r1x(x):
    if x == 0:
        # -(1/pi)(1/2)(psi(1)+psi(2))
        return -(mp.digamma(1) + mp.digamma(2)) / (2 * mp.pi)
    return (mp.bessely(1, x) - 2 / mp.pi * (mp.log(x / 2) * mp.besselj(1, x) - 1 / x)) / x


def pq(nu):
    def fp(s):
        u = (s + 1) / 2
        x = 8 / mp.sqrt(u)
        chi = x - (2 * nu + 1) * mp.pi / 4
        a = mp.sqrt(mp.pi * x / 2)
        return a * (mp.besselj(nu, x) * mp.cos(chi) + mp.bessely(nu, x) * mp.sin(chi))

    def fq(s):
        u = (s + 1) / 2
        x = 8 / mp.sqrt(u)
        z = 8 / x
        chi = x - (2 * nu + 1) * mp.pi / 4
        a = mp.sqrt(mp.pi * x / 2)
        return a * (-mp.besselj(nu, x) * mp.sin(chi) + mp.bessely(nu, x) * mp.cos(chi)) / z
    return fp, fq


def emit(name, coef):
    body = ",\n".join("    %s" % mp.nstr(c, 20, min_fixed=-1, max_fixed=-1) for c in coef)
    return "static constexpr int %s_N = %d;\nMG_BESSEL_TABLE double %s[%d] = {\n%s\n};\n" % (
        name, len(coef), name, len(coef), body)


def main():
    fits = [
        ("BJ0", trim(cheb_fit(small(j0)))),
        ("BR0", trim(cheb_fit(small(r0)))),
        ("BJ1", trim(cheb_fit(small(j1x)))),
        ("BR1", trim(cheb_fit(small(r1x)))),
    ]
    p0, q0 = pq(0)
    p1, q1 = pq(1)
    fits += [
        ("BP0", trim(cheb_fit(p0))),
        ("BQ0", trim(cheb_fit(q0))),
        ("BP1", trim(cheb_fit(p1))),
        ("BQ1", trim(cheb_fit(q1))),
    ]
    print("// Generated by gen_bessel.py (mpmath Chebyshev interpolation, 50 digits). Do not edit.")
    print("#pragma once")
    print("#ifndef MG_BESSEL_TABLE")
    print("#define MG_BESSEL_TABLE static constexpr")
    print("#endif")
    for name, c in fits:
        print(emit(name, c))


if __name__ == "__main__":
    main()
