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
next float up, 0x179abe16, is caught."""
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
