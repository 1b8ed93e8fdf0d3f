"""CPU replay of the resident migrate kernel's array indices (no GPU).

tests/bounds_replay.cpp walks one launch's grid as migrate_tile does (tiles,
waves, shot groups, receiver chunks or columns, every shot of a chunk's range,
every receiver of a chunk) through the kernel's own index functions
(kirch_datuming_amd/csrc/kmig_index.h) and checks each index against the size
the library allocates for that array.  Geometries: the shot-range split of
tests/test_gpu_split.py (the run where a NaN once appeared), the small workload
with 8x8 tiles and 16-receiver chunks, receiver columns (cmp=1), and the salt
workload's 16 shot groups at its edge tiles.  A planted mistake, the tap-row
addressing without the clamp of padding receivers to the shot's last trace
(the addressing that once faulted on the GPU), must be caught."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("replay") / "bounds_replay")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "kirch_datuming_amd", "csrc"),
                    os.path.join(ROOT, "tests", "bounds_replay.cpp"), "-o", exe], check=True)

    def run(*args):
        r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                           check=True)
        out = {}
        for line in r.stdout.split("\n"):
            f = line.split()
            if len(f) == 3:
                out[f[0]] = (int(f[1]), int(f[2]))
            elif len(f) == 2:
                out[f[0]] = int(f[1])
        return out
    return run


# nz nx TZ TX R ns nh nt sb se nsplit cols unclamped tile_step
CASES = {
    "small shots 5..45, 5 groups, 16x8 R8": (201, 301, 16, 8, 8, 61, 201, 1001, 5, 45, 5, 0, 0, 1),
    "small all shots, 7 groups, 8x8 R16": (201, 301, 8, 8, 16, 61, 201, 1001, 0, 61, 7, 0, 0, 1),
    "small all shots, 32x4 R8": (201, 301, 32, 4, 8, 61, 201, 1001, 0, 61, 7, 0, 0, 1),
    "cmp=1 columns, 3 groups, 16x8 R8": (60, 120, 16, 8, 8, 30, 24, 500, 0, 30, 3, 1, 0, 1),
    "cmp=1 columns, shots 3..29, 8x8 R16": (60, 120, 8, 8, 16, 30, 24, 500, 3, 29, 2, 1, 0, 1),
    "salt 16 groups, edge tiles": (400, 1000, 16, 8, 8, 768, 401, 2001, 0, 768, 16, 0, 0, 3124),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_every_index_stays_inside_its_array(replay, name):
    out = replay(*CASES[name])
    for arr, val in out.items():
        if arr != "violations" and val[0] > 0:
            assert val[0] <= val[1], (name, arr, val)
    assert out["violations"] == 0
    # the replay did reach the ends of the arrays it guards
    assert out["pairs"][0] > 0 and out["stup"][0] > 0


def test_replay_catches_the_unclamped_tap_rows(replay):
    """Without the clamp, padding receivers of the last chunk of the last shot
    (201 receivers in chunks of 8) read rows past the tap-pair array."""
    args = list(CASES["small all shots, 7 groups, 8x8 R16"])
    args[12] = 1
    out = replay(*args)
    assert out["violations"] > 0
    assert out["pairs"][0] > out["pairs"][1]
