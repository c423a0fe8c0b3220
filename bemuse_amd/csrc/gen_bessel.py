"""Generates bem_bessel_coeffs.h: the constants of the device Bessel J0, J1,
Y0, Y1 evaluator in bem_device.hpp (run: python gen_bessel.py > bem_bessel_coeffs.h).

Method (the Hankel modulus form, e.g. Abramowitz & Stegun 9.2.5-9.2.6):
for x >= 3 and chi_nu = x - (2 nu + 1) pi / 4,
    J_nu(x) = sqrt(2 / (pi x)) (P_nu(x) cos chi_nu - Q_nu(x) sin chi_nu)
    Y_nu(x) = sqrt(2 / (pi x)) (P_nu(x) sin chi_nu + Q_nu(x) cos chi_nu)
with P_nu, Q_nu smooth and non-oscillatory. Each of P0, Q0, P1, Q1 is
interpolated at the Chebyshev points of t = 2 (3 / x) - 1 in [-1, 1]
(degree 23; interpolation error below 1e-18 over x >= 3) from 50-digit
values of J and Y (mpmath). The interpolant is then rewritten exactly (in
50 digits) in powers of t and the device evaluates it by Horner's rule, one
fused multiply-add per term instead of Clenshaw's two operations. The power
form is well conditioned here: the sum of |a_j| is within 1% of |a_0| for
P0, P1 and twice |a_0| for Q0, Q1, so the rounding of the coefficients and of
Horner's rule stays within 2e-16 (measured against mpmath over x in
[3, 1e5]). The FP32 option takes the degree-10 truncation of the same
Chebyshev series, rewritten in powers of t the same way. For x <= 3 only J0, J1 are needed (the Struve
functions use Y0, Y1 only above 3, Kernels.cpp:35-116): their power series
sum_k (-x^2/4)^k / (k! (k + nu)!) truncated after 16 terms (last term below
1e-19). sin and cos of the reduced argument use their Taylor series
(degree 17 and 18, remainder below 1e-19 on |r| <= pi/4), and the reduction
x - n pi/2 uses pi/2 split into two 32-bit pieces and a remainder, so n
times each piece is exact for n < 2^21.
"""
import mpmath as mp

mp.mp.dps = 50
DEG = 23
DEG_F32 = 10
NTAYLOR = 16


def hankel_pq(x, nu):
    j, y = mp.besselj(nu, x), mp.bessely(nu, x)
    chi = x - (2 * nu + 1) * mp.pi / 4
    s = mp.sqrt(mp.pi * x / 2)
    return s * (j * mp.cos(chi) + y * mp.sin(chi)), s * (y * mp.cos(chi) - j * mp.sin(chi))


def chebyshev_coeffs(f, n):
    """Coefficients c_0..c_n of the degree-n interpolant of f(t) at the
    Chebyshev points of the first kind (exact-node DCT in 50-digit arithmetic)."""
    th = [mp.pi * (k + mp.mpf(1) / 2) / (n + 1) for k in range(n + 1)]
    fv = [f(mp.cos(t)) for t in th]
    c = [2 * mp.fsum(fv[i] * mp.cos(j * th[i]) for i in range(n + 1)) / (n + 1) for j in range(n + 1)]
    c[0] /= 2
    return c


def chebyshev_to_power(c):
    """a_0..a_n with sum_k c_k T_k(t) = sum_j a_j t^j (exact, from the
    recurrence T_k = 2 t T_{k-1} - T_{k-2})."""
    T = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
    for k in range(2, len(c)):
        a = [mp.mpf(0)] + [2 * v for v in T[k - 1]]
        b = T[k - 2] + [mp.mpf(0)] * (len(a) - len(T[k - 2]))
        T.append([x - y for x, y in zip(a, b)])
    a = [mp.mpf(0)] * len(c)
    for k in range(len(c)):
        for j, v in enumerate(T[k]):
            a[j] += c[k] * v
    return a


def t_to_x(t):
    u = (t + 1) / 2            # u = 3 / x in [0, 1]
    return 3 / u


def pq_of_t(nu, which):
    def f(t):
        if t <= -1:
            return mp.mpf(1) if which == 0 else mp.mpf(0)   # x -> infinity
        return hankel_pq(t_to_x(t), nu)[which]
    return f


def split_32(v):
    """Leading 32 significant bits of v (as an exactly representable double)."""
    e = mp.floor(mp.log(abs(v), 2))
    scale = mp.mpf(2) ** (31 - e)
    return mp.floor(v * scale) / scale


def lit(v):
    return repr(float(v))


def main():
    out = ["// Generated by gen_bessel.py (mpmath " + mp.__version__ + "); do not edit.",
           "// Constants of bessel01() in bem_device.hpp; the method is described in the generator.",
           "#pragma once", "", "namespace bemuse {", "namespace dev {", ""]
    for name, nu, which in (("kBesP0", 0, 0), ("kBesQ0", 0, 1), ("kBesP1", 1, 0), ("kBesQ1", 1, 1)):
        c = chebyshev_coeffs(pq_of_t(nu, which), DEG)
        for tab, deg in ((name, DEG), (name + "F", DEG_F32)):
            a = chebyshev_to_power(c[:deg + 1])
            out.append(f"// {'PQ'[which]}{nu}, x >= 3: degree-{deg} Chebyshev interpolant in powers of t = 6 / x - 1")
            out.append(f"constexpr double {tab}[{deg + 1}] = {{")
            for i in range(0, deg + 1, 3):
                out.append("    " + ", ".join(lit(v) for v in a[i:i + 3]) + ",")
            out.append("};")
    fac = mp.factorial
    j0 = [1 / (fac(k) * fac(k)) for k in range(NTAYLOR)]
    j1 = [1 / (fac(k) * fac(k + 1)) for k in range(NTAYLOR)]
    out.append("// power series of J0, J1 in -x^2/4: 1 / (k! k!), 1 / (k! (k + 1)!)")
    out.append(f"constexpr double kBesJ0S[{NTAYLOR}] = {{" + ", ".join(lit(v) for v in j0) + "};")
    out.append(f"constexpr double kBesJ1S[{NTAYLOR}] = {{" + ", ".join(lit(v) for v in j1) + "};")
    sinc = [(-1) ** k / fac(2 * k + 1) for k in range(9)]   # r^1 .. r^17
    cosc = [(-1) ** k / fac(2 * k) for k in range(10)]      # r^0 .. r^18
    out.append("// Taylor coefficients of sin r / r and cos r in r^2")
    out.append("constexpr double kSinT[9] = {" + ", ".join(lit(v) for v in sinc) + "};")
    out.append("constexpr double kCosT[10] = {" + ", ".join(lit(v) for v in cosc) + "};")
    half_pi = mp.pi / 2
    c1 = split_32(half_pi)
    c2 = split_32(half_pi - c1)
    c3 = half_pi - c1 - c2
    out.append("// pi/2 = kPio2A + kPio2B + kPio2C (A, B with 32 significant bits)")
    out.append(f"constexpr double kPio2A = {lit(c1)}, kPio2B = {lit(c2)}, kPio2C = {lit(c3)};")
    out.append(f"constexpr double kTwoOverPi = {lit(2 / mp.pi)};")
    out.append(f"constexpr double kTwoOver3Pi = {lit(2 / (3 * mp.pi))};")
    out.append(f"constexpr double kSqrtHalf = {lit(mp.sqrt(mp.mpf(1) / 2))};")
    out += ["", "}  // namespace dev", "}  // namespace bemuse", ""]
    print("\n".join(out))


if __name__ == "__main__":
    main()
