"""Generate the polynomial coefficients of the kernel's erfc (gmle_erf.hpp).

The kernels work in the scaled variable v = k u, k = sqrt(log2 e), so that the
Gaussian of every pixel edge is exp(-u^2) = exp2(-v^2) (one multiply and one
v_exp_f32).  erfc(|u|) = exp(-u^2) erfcx(|u|), and erfcx is evaluated as one
polynomial R of degree DEG in the variable
    t = 1 - 2 c / (|v| + c),   c = k C_U,
which maps |v| in [0, inf) onto t in [-1, 1) and turns erfcx's slow 1/u tail
into a smooth function of t (one add, one v_rcp_f32 and one FMA per edge):
    erfc(|u|) = exp2(-v^2) R(t)
R is fitted over |u| <= U_MAX = 10; beyond, exp2(-v^2) underflows Float32.
The kernels form pixel integrals from the edges' signed complements, so R's
relative accuracy carries over to the tails.

Fitting: least squares on Chebyshev nodes in u with Lawson's iteratively
reweighted refinement toward the minimax (equioscillating) relative error, in
float64; coefficients are then rounded to float32 and the float32 evaluation
(t with a correctly rounded reciprocal, and with the reciprocal moved by one
ulp either way, as v_rcp_f32 may round; FMA Horner) is measured on a dense
grid.

    python gaussmle.jl_amd/csrc/gen_erf_coeffs.py   # prints the C++ constants
"""
import numpy as np
from scipy.special import erfcx

K = float(np.sqrt(np.log2(np.e)))   # v = K u
U_MAX = 10.0
C_U = 2.5                           # c / k: where t = 0
DEG = 10


def t_of_v(v, c):
    return 1.0 - 2.0 * c / (np.abs(v) + c)


def lawson_fit(u, deg, iters=300):
    c = K * C_U
    t = t_of_v(K * u, c)
    y = erfcx(u)
    X = np.vander(t, deg + 1, increasing=True)
    W = 1.0 / np.abs(y)
    w = np.full(u.size, 1.0 / u.size)
    coef = None
    for _ in range(iters):
        sw = np.sqrt(w) * W
        coef, *_ = np.linalg.lstsq(X * sw[:, None], y * sw, rcond=None)
        r = np.abs((X @ coef - y) * W)
        w = w * r
        w /= w.sum()
    return coef


def eval32(coef32, v32, c32, ulp_shift=0, seed=0):
    """Float32 evaluation as the kernel does it: a = |v| + c, r = 1/a (optionally
    moved by one ulp, seeded), t = fma(-2c, r, 1), FMA Horner."""
    a = (np.abs(v32) + c32).astype(np.float32)
    r = (np.float32(1) / a).astype(np.float32)
    if ulp_shift:
        rng = np.random.default_rng(seed)
        to = np.where(rng.random(r.shape) < 0.5, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
        r = np.nextafter(r, to).astype(np.float32)
    t = (np.float64(-2 * c32) * r.astype(np.float64) + 1.0).astype(np.float32)
    p = np.full(v32.shape, coef32[-1], dtype=np.float32)
    for x in coef32[-2::-1]:
        p = (p.astype(np.float64) * t.astype(np.float64) + np.float64(x)).astype(np.float32)
    return p


def main():
    m = 30000
    nodes = (np.cos(np.pi * (np.arange(m) + 0.5) / m) + 1) / 2
    coef = lawson_fit(nodes * U_MAX, DEG).astype(np.float32)
    c32 = np.float32(K * C_U)
    v32 = (np.linspace(0, U_MAX, 400001) * K).astype(np.float32)
    ref = erfcx(v32.astype(np.float64) / K)
    e0 = np.abs(eval32(coef, v32, c32).astype(np.float64) / ref - 1).max()
    e1 = np.abs(eval32(coef, v32, c32, 1).astype(np.float64) / ref - 1).max()
    print(f"// erfcx: max relative error {e0:.3e} (reciprocal correctly rounded), {e1:.3e} (one ulp off)")
    print(f"constexpr float kErfK = {float(np.float32(K))!r}f;")
    print(f"constexpr float kErfC = {float(c32)!r}f;")
    print("constexpr float kErfR[%d] = {%s};" % (len(coef), ", ".join(f"{x!r}f" for x in map(float, coef))))


if __name__ == "__main__":
    main()
