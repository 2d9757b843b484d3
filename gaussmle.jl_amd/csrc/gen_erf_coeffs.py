"""Generate the polynomial coefficients of the kernel's erf (gmle_erf.hpp).

The kernels work in the scaled variable v = k u, k = sqrt(log2 e), so that the
Gaussian of every pixel edge is exp(-u^2) = exp2(-v^2) (one multiply and one
v_exp_f32).  erfc(|u|) = exp(-u^2) erfcx(|u|), and erfcx(v / k) is fitted as
one polynomial R of degree DEG in min(|v|, V_MAX) - C over [0, V_MAX]:
    erfc(|u|) = exp2(-v^2) R(min(|v|, V_MAX) - C)
erfcx is smooth there; V_MAX = k U_MAX with U_MAX = 5.6, where
erfc(5.6) = 1.5e-15 lies below every tail a fit can see (N times a pixel
integral against a background).  The kernels form pixel integrals from the
edges' signed complements, so R's relative accuracy carries over to the
tails.

Fitting: least squares on Chebyshev nodes with Lawson's iteratively reweighted
refinement toward the minimax (equioscillating) relative error, in float64;
coefficients are then rounded to float32 and the float32 FMA-Horner error is
measured on a dense grid.

    python gaussmle.jl_amd/csrc/gen_erf_coeffs.py   # prints the C++ constants
"""
import numpy as np
from scipy.special import erf, erfcx

K = float(np.sqrt(np.log2(np.e)))   # v = K u
U_MAX = 5.6
V_MAX = K * U_MAX
DEG = 16
C = V_MAX / 2


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
    v = nodes * V_MAX
    f = lambda x: erfcx(x / K)  # noqa: E731
    c = lawson_fit(f, v, DEG, C).astype(np.float32)
    vmax32, c32 = np.float32(V_MAX), np.float32(C)
    grid = np.linspace(0, V_MAX, 400001).astype(np.float32)
    r = horner32(c, (np.minimum(grid, vmax32) - c32).astype(np.float32))
    e_rel = np.abs(r.astype(np.float64) / f(grid.astype(np.float64)) - 1).max()
    ug = grid.astype(np.float64) / K
    g32 = np.exp2(-(grid.astype(np.float64) ** 2)).astype(np.float32)
    ec = (g32.astype(np.float64) * r.astype(np.float64)).astype(np.float32)
    v_erf = (np.float32(1) - ec).astype(np.float32)
    e_abs = np.abs(v_erf.astype(np.float64) - erf(ug)).max()
    print(f"// erfcx: max relative error {e_rel:.3e}; erf = 1 - erfc: max absolute error {e_abs:.3e}")
    print(f"constexpr float kErfK = {float(np.float32(K))!r}f;")
    print(f"constexpr float kErfVMax = {float(vmax32)!r}f;")
    print(f"constexpr float kErfC = {float(c32)!r}f;")
    print("constexpr float kErfR[%d] = {%s};" % (len(c), ", ".join(f"{x!r}f" for x in map(float, c))))


if __name__ == "__main__":
    main()
