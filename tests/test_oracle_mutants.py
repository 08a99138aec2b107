"""Mutation check of the oracle's closed-form pins (tests/test_oracle_closed_forms.py).

Each case plants one mistake in a copy of oracle/tc_oracle.c (a swapped
neighbour, a dropped factor, a wrong viscosity at a wall, a permuted stress
component), builds it into a temporary library, and runs the pin meant to catch
it against that library in a fresh interpreter (TC_ORACLE_LIB): the pin must
fail with an AssertionError. The unmodified source, built the same way, must
pass every pin. The mistakes are planted on the CPU only; nothing here runs the
engine.
"""
import concurrent.futures
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "tc_oracle.c")

# (case, pin that must fail, text in tc_oracle.c, planted replacement)
MUTANTS = [
    ("sxy_190 averaged at the geometric edge (the :598 average 'fixed')", "naive_linear_viscosity",
     "mu(YL_(in)) + mu(XR_(YR_(in))));", "mu(YL_(in)) + mu(XR_(YL_(in))));"),
    ("NAIVE wall stress of u with the total viscosity", "staggered_viscous_wall_rows",
     "double sxz_909 = (iz == 0) ? (2 * u_(in) * inv_hz * visc)",
     "double sxz_909 = (iz == 0) ? (2 * u_(in) * inv_hz * mu(in))"),
    ("NAIVE szz below the face with the cell's own viscosity", "staggered_viscous_wall_rows",
     "double szz_008 = 2 * (w_(in) - (iz == 0 ? 0.0 : w_(ZL_(in)))) * mu(ZL_(in)) / DZ[iz - 1];",
     "double szz_008 = 2 * (w_(in) - (iz == 0 ? 0.0 : w_(ZL_(in)))) * mu(in) / DZ[iz - 1];"),
    ("CROSS top-wall distance DZ instead of DZ/2", "staggered_viscous_wall_rows",
     "        double inv_hz_up = 2.0 / (DZ[iz]);", "        double inv_hz_up = 1.0 / (DZ[iz]);"),
    ("Galerkin element to the right instead of the left", "galerkin_element_viscosity",
     "    if (!ox) in = XL_(in);", "    if (!ox) in = XR_(in);"),
    ("Galerkin element viscosity with y and z offsets swapped", "galerkin_element_viscosity",
     "nu_t[element_index(s, in, ox, oy, oz)]", "nu_t[element_index(s, in, ox, oz, oy)]"),
    ("collocated wall node of full height", "collocated_integrals",
     "    if (iz == 0) return 0.5 * s->DZ[0];", "    if (iz == 0) return s->DZ[0];"),
    ("collocated viscous CFL factor", "collocated_cfl",
     "            fac = 4 * (s->P.visc_type + 1);", "            fac = 4 * (s->P.visc_type + 2);"),
    ("node viscosity as a quarter of the element sum", "collocated_cfl",
     "    return 0.125 * (n[i] + n[XL_(i)]", "    return 0.25 * (n[i] + n[XL_(i)]"),
    ("profile XZ column holding v w", "collocated_profiles",
     "sm[6] += u * v; sm[7] += u * w; sm[8] += v * w;", "sm[6] += u * v; sm[7] += v * w; sm[8] += v * w;"),
    ("staggered w volume from the cell above only", "staggered_integrals_and_cfl",
     "c += (i < NXY ? 0.0 : u[3 * i + 2] * 0.5 * (s->DZ[iz] + s->DZ[iz - 1])) * hx * hy;",
     "c += (i < NXY ? 0.0 : u[3 * i + 2] * 0.5 * (s->DZ[iz] + s->DZ[iz])) * hx * hy;"),
    ("staggered viscous CFL without nu_t", "staggered_integrals_and_cfl",
     "            visc_coeff = s->nu_t ? s->P.visc + s->nu_t[i] : s->P.visc;",
     "            visc_coeff = s->nu_t ? s->P.visc : s->P.visc;"),
    ("staggered profile w not averaged to the cell centre", "staggered_profiles",
     "(iz == NZ - 1 ? 0.0 : vel[3 * (in + NXY) + 2])", "(iz == NZ - 1 ? 0.0 : vel[3 * (in) + 2])"),
    ("z derivative of tau with the coefficients of the other side", "collocated_stress_divergence",
     "const double cp = hm * hm / (hp * hm * (hp + hm))", "const double cp = hp * hp / (hp * hm * (hp + hm))"),
    ("d tau_zz / dz read from the YY component", "collocated_stress_divergence",
     "{{0, 3, 4}, {3, 1, 5}, {4, 5, 2}}", "{{0, 3, 4}, {3, 1, 5}, {4, 5, 1}}"),
]
PIN_NAMES = sorted({m[1] for m in MUTANTS})


def _build(tmpdir, name, old, new):
    src = open(SRC).read()
    if old is not None:
        assert src.count(old) == 1, f"planted text not found once: {old!r}"
        src = src.replace(old, new)
    c = os.path.join(tmpdir, f"{name}.c")
    so = os.path.join(tmpdir, f"lib{name}.so")
    with open(c, "w") as f:
        f.write(src)
    r = subprocess.run(["gcc", "-O1", "-std=gnu11", "-fPIC", "-fopenmp", "-shared", "-I", os.path.join(ROOT, "oracle"),
                        c, "-o", so, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return so


def _run_pins(lib, names):
    env = dict(os.environ, TC_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    return subprocess.run([sys.executable, "-m", "tests.test_oracle_closed_forms", *names], cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("oracle_mutants"))
    jobs = [("control", None, None)] + [(f"m{i}", m[2], m[3]) for i, m in enumerate(MUTANTS)]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        libs = list(ex.map(lambda j: _build(d, *j), jobs))
    return dict(zip([j[0] for j in jobs], libs))


def test_unmodified_oracle_passes_every_pin(builds):
    r = _run_pins(builds["control"], PIN_NAMES)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("i", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_its_pin(builds, i):
    case, pin, _, _ = MUTANTS[i]
    r = _run_pins(builds[f"m{i}"], [pin])
    assert r.returncode != 0, f"pin {pin} did not catch: {case}"
    assert "AssertionError" in r.stderr, r.stderr[-3000:]
