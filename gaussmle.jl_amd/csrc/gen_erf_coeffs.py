"""Generate the polynomial coefficients of the kernel's erf (gmle_erf.hpp).

erf(u), u >= 0, is evaluated in two pieces whose coefficients this script fits:
  * u <  T : erf(u) = c0 u + u s Q(s), s = u^2   Q of degree DEG_SMALL - 1, c0 = 2/sqrt(pi)
                                                 (the leading term is applied last in one FMA)
  * u >= T : erf(u) = 1 - exp(-u^2) * R(u - C)   R of degree DEG_LARGE, R ~ erfcx
erfcx(u) = exp(u^2) erfc(u) is smooth on [T, U_MAX]; the kernel already has
exp(-u^2) at every pixel edge (it is the Gaussian of the derivative tables), so
the large piece costs no extra transcendental.

Fitting: least squares on Chebyshev nodes with Lawson's iteratively reweighted
refinement toward the minimax (equioscillating) relative error, in float64;
coefficients are then rounded to float32 and the float32 FMA-Horner error is
measured on a dense grid.

    python gaussmle.jl_amd/csrc/gen_erf_coeffs.py   # prints the C++ constants
"""
import numpy as np
from scipy.special import erf, erfcx

T = 0.875
U_MAX = 3.92        # erf(u) rounds to 1.0f beyond this
DEG_SMALL = 6
DEG_LARGE = 12
C_LARGE = (T + U_MAX) / 2


def lawson_fit(f, x, deg, center, iters=200):
    X = np.vander(x - center, deg + 1, increasing=True)
    y = f(x)
    W = 1.0 / np.abs(y)
    w = np.full(x.size, 1.0 / x.size)
    coef = None
    for _ in range(iters):
        sw = np.sqrt(w) * W
        coef, *_ = np.linalg.lstsq(X * sw[:, None], y * sw, rcond=None)
        r = np.abs((X @ coef - y) * W)
        w = w * r
        w /= w.sum()
    return coef


def horner32(coef32, t32):
    """float32 Horner with fused multiply-add (one rounding per step)."""
    p = np.full(t32.shape, coef32[-1], dtype=np.float32)
    for a in coef32[-2::-1]:
        p = (p.astype(np.float64) * t32.astype(np.float64) + np.float64(a)).astype(np.float32)
    return p


def main():
    m = 20000
    nodes = (np.cos(np.pi * (np.arange(m) + 0.5) / m) + 1) / 2
    # small piece: erf(u)/u as a polynomial in s = u^2
    s = nodes * T * T
    f_small = lambda s: np.where(s > 0, erf(np.sqrt(s)) / np.sqrt(np.maximum(s, 1e-300)), 2 / np.sqrt(np.pi))
    c0 = 2 / np.sqrt(np.pi)
    f_q = lambda s: np.where(s > 1e-12, (f_small(s) - c0) / np.maximum(s, 1e-300), -c0 / 3)
    cq = lawson_fit(lambda s: f_q(s), s, DEG_SMALL - 1, 0.0).astype(np.float32)
    # large piece: erfcx(u) in (u - C)
    u = T + nodes * (U_MAX - T)
    cl = lawson_fit(erfcx, u, DEG_LARGE, C_LARGE).astype(np.float32)

    grid_s = np.linspace(1e-7, T, 400001).astype(np.float32)
    s32 = (grid_s * grid_s).astype(np.float32)
    q = horner32(cq, s32)
    us = (grid_s.astype(np.float64) * s32).astype(np.float32)                    # u * s
    tail = (us.astype(np.float64) * q.astype(np.float64)).astype(np.float32)     # (u s) Q(s)
    vs = (grid_s.astype(np.float64) * np.float64(np.float32(c0)) + tail).astype(np.float32)
    e_small = np.abs(vs.astype(np.float64) / erf(grid_s.astype(np.float64)) - 1).max()
    grid_l = np.linspace(T, U_MAX, 400001).astype(np.float32)
    vl = horner32(cl, (grid_l - np.float32(C_LARGE)).astype(np.float32))
    e_large = np.abs(vl.astype(np.float64) / erfcx(grid_l.astype(np.float64)) - 1).max()
    print(f"// small piece: max relative error of erf {e_small:.3e}")
    print(f"// large piece: max relative error of erfcx {e_large:.3e}")
    print(f"constexpr float kErfT = {T}f;")
    print(f"constexpr float kErfUMax = {U_MAX}f;")
    print(f"constexpr float kErfC = {float(np.float32(C_LARGE))!r}f;")
    print(f"constexpr float kErfC0 = {float(np.float32(c0))!r}f;")
    print("constexpr float kErfSmall[%d] = {%s};" % (len(cq), ", ".join(f"{c!r}f" for c in map(float, cq))))
    print("constexpr float kErfLarge[%d] = {%s};" % (len(cl), ", ".join(f"{c!r}f" for c in map(float, cl))))


if __name__ == "__main__":
    main()
