"""Generate bessel_ref_coeffs.h: the rational-approximation coefficients that
torch's CPU special functions use for J0, J1, Y0, Y1.

The reference computes its 2D Green's functions with torch.special.bessel_*
(Transmission2D.py:56-59, 588-591).  Those are torch's restatement of the
Cephes routines j0.c / j1.c (S. L. Moshier, Cephes Math Library, public
domain).  To match the reference's numbers, the kernels evaluate the same
rational forms, so this script copies the coefficient tables (data only) out
of the torch header installed with this Python (ATen/native/Math.h).  The
evaluation code is in bessel_ref.h.  The squared zeros that the small-x
branches factor out are computed here with mpmath.
Run: python gen_bessel_ref.py > bessel_ref_coeffs.h
"""
import os
import re

import mpmath as mp
import torch

FUNCS = ["bessel_j0_forward", "bessel_j1_forward", "bessel_y0_forward", "bessel_y1_forward"]


def tables(src, fn):
    start = src.index("inline C10_HOST_DEVICE T %s(T x)" % fn)
    end = src.index("} // %s(T x)" % fn, start)
    body = src[start:end]
    out = {}
    for m in re.finditer(r"static const T (\w+)\[\] = \{(.*?)\};", body, re.S):
        vals = [v.strip() for v in m.group(2).split(",") if v.strip()]
        out[m.group(1)] = vals
    return out


def main():
    hdr = os.path.join(os.path.dirname(torch.__file__), "include", "ATen", "native", "Math.h")
    src = open(hdr).read()
    print("// Generated by gen_bessel_ref.py from the Cephes coefficient tables as")
    print("// installed with torch %s (ATen/native/Math.h). Do not edit." % torch.__version__)
    print("#pragma once")
    for fn in FUNCS:
        name = fn.replace("bessel_", "").replace("_forward", "").upper()
        for tab, vals in tables(src, fn).items():
            print("static constexpr double REF_%s_%s[%d] = {%s};" % (name, tab, len(vals), ", ".join(vals)))
    mp.mp.dps = 40
    for nu in (0, 1):
        for k in (1, 2):
            z = mp.besseljzero(nu, k)
            print("static constexpr double REF_J%d_ZERO%d_SQ = %s;" % (nu, k, mp.nstr(z * z, 25)))


if __name__ == "__main__":
    main()
