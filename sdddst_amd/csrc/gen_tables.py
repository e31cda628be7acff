"""Generate libm_tables.h: the constant tables of the device libm.

Both tables are values of elementary functions at grid points, rounded to
double and split into a leading double plus a rounded remainder, computed here
with mpmath at 300 bits:

  exp2_tab[2k], exp2_tab[2k+1]  (k = 0..127):
      H = RN(2^(k/128)), T = RN((2^(k/128) - H) / H); stored as the bit
      patterns of T and of H minus k*2^45 (the exponent bits the caller adds
      back from the integer part of x*128/ln2).
  sincos_tab[4i .. 4i+3]        (i = 0..110):
      RN(sin(i/128)), RN(sin(i/128) - hi), RN(cos(i/128)), RN(cos(i/128) - hi).
  sdlibm_exp_2pik[k], sdlibm_exp_m2pik[k]  (k = 0..5):
      RN(exp(2 pi k)), RN(exp(-2 pi k)).

Run: python sdddst_amd/csrc/gen_tables.py > sdddst_amd/csrc/libm_tables.h
"""
import struct
import mpmath

mpmath.mp.prec = 300


def u64(d):
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def main():
    out = ["// Generated by gen_tables.py (mpmath, 300-bit).  Do not edit.", "#pragma once", "#include <stdint.h>", ""]
    out.append("// 2^(k/128) ~= H*(1+T): {bits(T), bits(H) - (k << 45)} for k = 0..127")
    out.append("SDLIBM_TABLE const uint64_t sdlibm_exp2_tab[256] = {")
    for k in range(128):
        v = mpmath.mpf(2) ** (mpmath.mpf(k) / 128)
        h = float(v)
        t = float((v - mpmath.mpf(h)) / mpmath.mpf(h))
        out.append("    0x%016xULL, 0x%016xULL," % (u64(t), (u64(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    out.append("};")
    out.append("")
    out.append("// {sin hi, sin lo, cos hi, cos lo} at i/128 for i = 0..110")
    out.append("SDLIBM_TABLE const double sdlibm_sincos_tab[444] = {")
    for i in range(111):
        x = mpmath.mpf(i) / 128
        s, c = mpmath.sin(x), mpmath.cos(x)
        sh, ch = float(s), float(c)
        sl, cl = float(s - mpmath.mpf(sh)), float(c - mpmath.mpf(ch))
        out.append("    %s, %s, %s, %s," % tuple(v.hex() for v in (sh, sl, ch, cl)))
    out.append("};")
    out.append("")
    out.append("// exp(+-2 pi k), k = 0..5")
    for name, sgn in (("sdlibm_exp_2pik", 1), ("sdlibm_exp_m2pik", -1)):
        vals = [float(mpmath.exp(sgn * 2 * mpmath.pi * k)) for k in range(6)]
        out.append("constexpr double %s[6] = {%s};" % (name, ", ".join(v.hex() for v in vals)))
    print("\n".join(out))


if __name__ == "__main__":
    main()
