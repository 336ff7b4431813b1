"""The oracle's pins catch plausible restatement mistakes (CPU only).

Each case compiles oracle/mst_oracle.c with one planted mistake (-DORC_PLANT=k, see the source) into a
temporary library, runs the pins that should catch it against that library in a fresh process
(MST_ORACLE_LIB), and requires them to fail.  A control build without a plant must pass the same
pins, so a failure is the plant's doing and not the build's."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "mst_oracle.c")
CFLAGS = ["-O2", "-std=c99", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared"]

PINS = "tests/test_oracle_pins.py"
KAT = "tests/test_oracle_kat.py"
VISC = "tests/test_oracle_viscous.py"
# planted mistake -> (what it is, the pins that must catch it)
PLANTS = {
    1: ("wall ghost: factor 2 dropped", [f"{PINS}::test_wall_ghost_on_slanted_walls",
                                         f"{PINS}::test_wall_mirror_fluxes_vanish",
                                         f"{PINS}::test_one_dimensional_reduction_on_a_strip"]),
    2: ("wall ghost: unnormalised area vector", [f"{PINS}::test_wall_ghost_on_slanted_walls",
                                                 f"{PINS}::test_wall_mirror_fluxes_vanish",
                                                 f"{PINS}::test_one_dimensional_reduction_on_a_strip"]),
    3: ("wall ghost: sign flipped", [f"{PINS}::test_wall_ghost_on_slanted_walls",
                                     f"{PINS}::test_wall_mirror_fluxes_vanish",
                                     f"{PINS}::test_one_dimensional_reduction_on_a_strip"]),
    4: ("3D kinetic energy over two components", [f"{PINS}::test_3d_residual_is_symmetric_under_axis_permutations"]),
    5: ("Roe shear eigenvector entry", [f"{KAT}::test_roe_eigenvectors_and_inverse"]),
    6: ("3D area vector: transposed operand", [f"{PINS}::test_3d_freestream_and_conservation"]),
    7: ("gather: c1 keeps c0's normal", [f"{PINS}::test_3d_freestream_and_conservation",
                                         f"{KAT}::test_freestream_preservation_and_conservation"]),
    8: ("Riemann sides swapped (downwind)", [f"{PINS}::test_one_dimensional_reduction_on_a_strip",
                                             f"{PINS}::test_3d_one_dimensional_reduction"]),
    9: ("viscous gradient: slip wall", [f"{VISC}::test_noslip_wall_closed_forms"]),
    10: ("viscous wall: inviscid mirror ghost", [f"{VISC}::test_noslip_wall_closed_forms"]),
}


def _build(tmp_path, plant):
    so = str(tmp_path / f"libmst_oracle_plant{plant}.so")
    subprocess.run(["gcc", *CFLAGS, f"-DORC_PLANT={plant}", "-o", so, SRC, "-lm"], check=True)
    return so


def _run(so, tests):
    env = dict(os.environ, MST_ORACLE_LIB=so, OMP_NUM_THREADS="2")
    # every listed test runs (no -x); the exit status is 1 when any of them fails
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "--tb=no", *tests],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


def _failed(out):
    return sorted({ln.split(" ")[1].split("[")[0] for ln in out.splitlines() if ln.startswith("FAILED ")})


def test_control_build_passes_every_pin(tmp_path):
    tests = sorted({t for _, ts in PLANTS.values() for t in ts})
    r = _run(_build(tmp_path, 0), tests)
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_caught(tmp_path, plant):
    what, tests = PLANTS[plant]
    r = _run(_build(tmp_path, plant), tests + ["-rf"])
    assert r.returncode == 1, f"plant {plant} ({what}) not caught:\n{r.stdout[-2000:]}"
    caught = _failed(r.stdout)
    # every listed pin catches it, not just one of them
    assert caught == sorted(tests), (what, caught)
