"""Coefficients of the polynomial kernels of lt_math.h (atan_lt, log_lt), derived here from the
mathematics with mpmath (Chebyshev interpolation at 50 significant digits), printed as C
hexadecimal floating-point literals. Not part of the build: the printed tables are in lt_math.h;
run `python gen_poly.py` to reproduce them.

  atan(t) = t + t^3 P(t^2),            |t| <= tan(pi/8),   P of degree 10 (FP64), 4 (FP32: kAtanP32)
  log(m)  = 2 s + 2 s^3 R(s^2),        s = (m - 1)/(m + 1), m in [sqrt(1/2), sqrt(2)),  R of degree 6

The interpolation error (printed) is far below the rounding of a double."""
import mpmath as mp

mp.mp.dps = 50


def fit(f, b, deg):
    poly, err = mp.chebyfit(f, [0, b], deg + 1, error=True)
    return list(reversed(poly)), err  # ascending powers


def P(z):
    if z == 0:
        return mp.mpf(-1) / 3
    r = mp.sqrt(z)
    return (mp.atan(r) / r - 1) / z


def R(w):
    if w == 0:
        return mp.mpf(1) / 3
    r = mp.sqrt(w)
    return (mp.atanh(r) / r - 1) / w


def main():
    a = mp.sqrt(2) - 1
    s = (mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)
    for name, f, b, deg in (("kAtanP", P, a * a, 10), ("kLogR", R, s * s, 6)):
        c, err = fit(f, b, deg)
        print(f"// {name}: degree {deg} on [0, {mp.nstr(b, 6)}], interpolation error {mp.nstr(err, 3)}")
        print(f"constexpr double {name}[{deg + 1}] = {{")
        for v in c:
            print(f"    {float(v).hex()},  // {mp.nstr(v, 20)}")
        print("};")
    # FP32 table: the coefficients rounded to float (struct round trip)
    import struct
    c, err = fit(P, a * a, 4)
    print(f"// kAtanP32: degree 4 on [0, {mp.nstr(a * a, 6)}], error in atan {mp.nstr(err * a ** 3, 3)}")
    print("constexpr float kAtanP32[5] = {")
    for v in c:
        f32 = struct.unpack("f", struct.pack("f", float(v)))[0]
        print(f"    {float(f32).hex()}f,  // {mp.nstr(v, 12)}")
    print("};")


if __name__ == "__main__":
    main()
