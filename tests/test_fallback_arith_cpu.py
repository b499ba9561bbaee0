"""The arithmetic of the tiled kernels' fall-back scatter, restated on the host: each particle's
FP32 contributions are rounded to its grain 2^(ex - 26) (bound < 2^ex) by (x + M) - M in FP64 with
M = 1.5 * 2^(ex + 26), and the rounded values are summed in FP64.  The sums must not depend on the
order of the adds while the bounds at a node lie within the stated range, and the rounding error
must stay within half the grain, under 2^-26 of the particle's bound."""
import numpy as np


def magic(bound):
    ex = np.frexp(np.float64(np.float32(bound)))[1]   # bound < 2^ex
    return np.ldexp(1.5, int(ex) + 26)


def grain_round(x, M):
    return (np.float64(np.float32(x)) + M) - M


def test_rounding_error_and_grain():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        bound = np.float32(10.0 ** rng.uniform(-12, 6))
        M = magic(bound)
        x = np.float32(rng.uniform(-1, 1) * bound)
        q = grain_round(x, M)
        ex = int(np.frexp(np.float64(bound))[1])
        g = 2.0 ** (ex - 26)                     # the spacing of M's binade [2^(ex+26), 2^(ex+27))
        assert np.spacing(M) == g
        assert abs(q - np.float64(x)) <= 0.5 * g
        assert q / g == np.rint(q / g)           # a multiple of the grain
        assert 0.5 * g <= 2.0 ** -26 * float(bound)


def test_sums_do_not_depend_on_order():
    """Contributions of particles whose bounds span 2^12 (inside the exact range for 64 of them)
    give the same FP64 sum in 200 random orders; with bounds spanning 2^60 the sum is still
    within FP64 rounding of the exact one."""
    rng = np.random.default_rng(1)
    for span in (12.0, 60.0):
        bounds = np.float32(2.0 ** rng.uniform(-span / 2, span / 2, size=64))
        vals = np.array([grain_round(np.float32(rng.uniform(-1, 1) * b), magic(b)) for b in bounds])
        sums = set()
        for _ in range(200):
            s = np.float64(0.0)
            for v in rng.permutation(vals):
                s = s + v
            sums.add(float(s))
        exact = float(np.sum(np.array(vals, dtype=np.longdouble)))
        if span == 12.0:
            assert len(sums) == 1
            assert list(sums)[0] == exact
        else:
            assert all(abs(s - exact) <= 64 * np.spacing(abs(exact) + max(abs(vals))) for s in sums)
