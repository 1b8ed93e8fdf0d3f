"""The box smoothing of the fused tuple pass divides its interior sums by 2r + 1
as a reciprocal product with one fused correction (div_box,
kirch_datuming_amd/csrc/kmig_prep.hip); the packed tuples stay bitwise those of
the separate passes only if that equals IEEE division.  tests/div_exact.c checks
it over every finite float (the non-negative ones; both sides are symmetric in
sign) for the four divisors of smoothing radii 1 to 4; -0.0, the one exception,
does not arise in a box sum, which starts from +0.0.  A planted wrong divisor, 6, for which the product is not exact,
must be caught.

The same pass also tests whether a smoothed gradient vanishes as |g|^2 <= S
instead of sqrtf(|g|^2) <= 1e-12 (grad_vanishes); tests/sqrt_cut.c checks the
two forms agree on every non-negative float for S = 0x179abe15, and that the
next float up, 0x179abe16, is caught.

The fused pass also divides by divisors fixed per launch -- the interior
finite-difference spans (2 r) dx, (2 r) dz and pi in the aperture angle -- as
the float of x times the divisor's double reciprocal (div_fixed), equal to the
IEEE quotient for |x| >= 2^-60 by a rounding argument; tests/recip_exact.c
checks every such float for the divisors of the workloads here and others,
and that the float reciprocal product (a plausible slip) is caught."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def div_exact(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("div") / "div_exact")
    subprocess.run(["gcc", "-O2", "-march=native", "-fopenmp", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "div_exact.c"), "-o", exe, "-lm"], check=True)

    def run(*ds):
        r = subprocess.run([exe] + [str(d) for d in ds], capture_output=True, text=True,
                           check=True, env=dict(os.environ, OMP_NUM_THREADS=str(min(16, os.cpu_count() or 1))))
        return {int(f.split()[0][2:]): int(f.split()[2]) for f in r.stdout.strip().split("\n")}
    return run


def test_reciprocal_product_is_ieee_division_for_box_counts(div_exact):
    assert div_exact(3, 5, 7, 9) == {3: 0, 5: 0, 7: 0, 9: 0}


def test_check_catches_a_divisor_where_the_product_is_not_exact(div_exact):
    assert div_exact(6)[6] > 0


@pytest.fixture(scope="module")
def sqrt_cut(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sqrt") / "sqrt_cut")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "sqrt_cut.c"), "-o", exe, "-lm"], check=True)

    def run(*cuts):
        r = subprocess.run([exe] + list(cuts), capture_output=True, text=True, check=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(min(16, os.cpu_count() or 1))))
        return {f.split()[0][4:]: int(f.split()[2]) for f in r.stdout.strip().split("\n")}
    return run


def test_root_free_gradient_test_is_the_rooted_one(sqrt_cut):
    assert sqrt_cut("179abe15", "179abe16") == {"179abe15": 0, "179abe16": 1}


@pytest.fixture(scope="module")
def recip_exact(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("recip") / "recip_exact")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "recip_exact.c"), "-o", exe, "-lm"], check=True)

    def run(*ds):
        r = subprocess.run([exe] + list(ds), capture_output=True, text=True, check=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(min(16, os.cpu_count() or 1))))
        return {f.split()[0][2:]: int(f.split()[2]) for f in r.stdout.strip().split("\n")}
    return run


def test_double_reciprocal_product_is_ieee_division(recip_exact):
    # (2 r) d for radii 1..4 and the grid spacings 0.01, 0.02, 0.015; pi as a
    # float; and a few others
    import numpy as np
    ds = sorted({repr(float(np.float32(2 * r) * np.float32(h)))
                 for r in (1, 2, 3, 4) for h in (0.01, 0.02, 0.015)})
    ds += [repr(float(np.float32(np.pi))), "7.0", "0.001", "123.456"]
    assert all(v == 0 for v in recip_exact(*ds).values())


def test_check_catches_the_float_reciprocal(recip_exact):
    assert recip_exact("f:0.04")["f:0.04"] > 0
