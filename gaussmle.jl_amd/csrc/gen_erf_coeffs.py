"""Generate the polynomial coefficients of the kernel's erf (gmle_erf.hpp).

erf(u), u >= 0, is evaluated as one piece over [0, U_MAX]:
    erf(u) = 1 - exp(-u^2) * R(min(u, U_MAX) - C),   R of degree DEG ~ erfcx
erfcx(u) = exp(u^2) erfc(u) is smooth on [0, U_MAX]; the kernel already has
exp(-u^2) at every pixel edge (it is the Gaussian of the derivative tables), so
erf costs one polynomial and no transcendental of its own.  Beyond U_MAX,
exp(-u^2) R < 3e-8 and 1 - exp(-u^2) R rounds to 1.0f by itself.  Near u = 0
the form is accurate in absolute terms (~1e-7) rather than relative ones,
which is what the pixel integrals Phi = (erf(b) - erf(a)) / 2 consume.

Fitting: least squares on Chebyshev nodes with Lawson's iteratively reweighted
refinement toward the minimax (equioscillating) relative error, in float64;
coefficients are then rounded to float32 and the float32 FMA-Horner error is
measured on a dense grid.

    python gaussmle.jl_amd/csrc/gen_erf_coeffs.py   # prints the C++ constants
"""
import numpy as np
from scipy.special import erf, erfcx

U_MAX = 3.92        # erf(u) rounds to 1.0f beyond this
DEG = 15
C = U_MAX / 2


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
    u = nodes * U_MAX
    c = lawson_fit(erfcx, u, DEG, C).astype(np.float32)
    grid = np.linspace(0, U_MAX, 400001).astype(np.float32)
    r = horner32(c, (grid - np.float32(C)).astype(np.float32))
    e_rel = np.abs(r.astype(np.float64) / erfcx(grid.astype(np.float64)) - 1).max()
    g32 = np.exp(-(grid.astype(np.float64) ** 2)).astype(np.float32)
    v = (np.float32(1) - (g32.astype(np.float64) * r.astype(np.float64)).astype(np.float32)).astype(np.float32)
    e_abs = np.abs(v.astype(np.float64) - erf(grid.astype(np.float64))).max()
    print(f"// erfcx: max relative error {e_rel:.3e}; erf: max absolute error {e_abs:.3e}")
    print(f"constexpr float kErfUMax = {U_MAX}f;")
    print(f"constexpr float kErfC = {float(np.float32(C))!r}f;")
    print("constexpr float kErfR[%d] = {%s};" % (len(c), ", ".join(f"{x!r}f" for x in map(float, c))))


if __name__ == "__main__":
    main()
