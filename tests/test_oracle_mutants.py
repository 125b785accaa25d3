"""Mutation check of the oracle's pins: plausible mistakes planted in a scratch copy of
oracle/oracle.cpp (a dropped factor, a wrong sign, a full step instead of a half step, a
missing boundary condition, ...) must each make the closed-form pins of test_oracle_pins.py
fail. Each mutant is compiled on the host into a temporary directory and loaded by a pytest
subprocess through MFC_ORACLE_LIB; the oracle in the repository is never modified."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "oracle", "oracle.cpp")
PINS = os.path.join(REPO, "tests", "test_oracle_pins.py")

# (name, text in oracle.cpp, replacement); each text occurs exactly once
MUTANTS = [
    ("viscosity_sign",
     "vt.x += -mass * piij * w_x;\n              vt.y += -mass * piij * w_y;",
     "vt.x += mass * piij * w_x;\n              vt.y += mass * piij * w_y;"),
    ("artificial_stress_f1",
     "            fab *= fab;\n            fab *= fab;\n", ""),
    ("artificial_stress_f2",
     "            fab *= fab;\n            fab *= fab;\n", "            fab *= fab;\n"),
    ("eos_c0_not_squared",
     "s.p[i] = c0 * c0 * (s.rho[i] - s.rho0);", "s.p[i] = c0 * (s.rho[i] - s.rho0);"),
    ("xsph_twice",
     "xt.x += -eps * w * mass / rhoij * vijx;\n            xt.y += -eps * w * mass / rhoij * vijy;",
     "xt.x += -2 * eps * w * mass / rhoij * vijx;\n            xt.y += -2 * eps * w * mass / rhoij * vijy;"),
    ("fric_heat_twice",
     "s.T[i] += s.eta_heat * dt * fmag * vmag / (s.cp * s.mass);",
     "s.T[i] += 2 * s.eta_heat * dt * fmag * vmag / (s.cp * s.mass);"),
    ("fric_heat_tangential_projection",
     "V2 vr = v2(vd.x - vdn, vd.y - vdn);", "V2 vr = v2(vd.x - vdn * nrm.x, vd.y - vdn * nrm.y);"),
    ("predict_T_full_step",
     "s.T[i] = s.T_init_arr[i] + 0.5 * dt * s.T_t[i];", "s.T[i] = s.T_init_arr[i] + dt * s.T_t[i];"),
    ("predict_S_full_step",
     "s.S[i].x = s.S_init[i].x + 0.5 * dt * s.S_t[i].x;", "s.S[i].x = s.S_init[i].x + dt * s.S_t[i].x;"),
    ("predict_rho_full_step",
     "s.rho[i] = s.rho_init[i] + 0.5 * dt * s.rho_t[i];", "s.rho[i] = s.rho_init[i] + dt * s.rho_t[i];"),
    ("predict_pos_full_step",
     "s.pos[i].x = s.pos_init[i].x + 0.5 * dt * s.pos_t[i].x;", "s.pos[i].x = s.pos_init[i].x + dt * s.pos_t[i].x;"),
    ("predict_vel_skipped",
     "s.vel[i].x = s.vel_init[i].x + 0.5 * dt * s.vel_t[i].x;", "s.vel[i].x = s.vel_init[i].x;"),
    ("fixed_vel_t_not_zeroed",
     "s.pos_t[i] = v2(0., 0.); s.vel_t[i] = v2(0., 0.);", "s.pos_t[i] = v2(0., 0.);"),
    ("thermal_bc_skips_tool",
     "if (s.fixed[i] == 1.) s.T[i] = s.T_init;", "if (s.fixed[i] == 1. && s.tool_p[i] == 0.) s.T[i] = s.T_init;"),
    ("stress_divergence_drops_neighbour",
     "Sd.x += mass * ((Si.x - prsi) * rhoi21 + (Sj.x - prsj) * rhoj21 + Rx) * w_x;",
     "Sd.x += mass * ((Si.x - prsi) * rhoi21 + Rx) * w_x;"),
]


def build_mutant(name, old, new, tmp):
    src = open(SRC).read()
    if old is not None:
        assert src.count(old) == 1, f"{name}: the mutated text must occur once in oracle.cpp"
        src = src.replace(old, new)
    path = os.path.join(tmp, "oracle.cpp")
    with open(path, "w") as f:
        f.write(src)
    lib = os.path.join(tmp, "liboracle.so")
    # the same flags as oracle/Makefile; -iquote resolves oracle.cpp's relative includes
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    "-iquote", os.path.join(REPO, "oracle"), "-shared", "-o", lib, path], check=True)
    return lib


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_the_mutant(name, old, new, tmp_path):
    lib = build_mutant(name, old, new, str(tmp_path))
    env = dict(os.environ, MFC_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", PINS],
                       env=env, cwd=REPO, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, f"{name}: the pins did not fail (rc {r.returncode}):\n{r.stdout[-2000:]}"


def test_the_unmutated_copy_passes(tmp_path):
    """The harness itself: the same build of the unmodified source passes the pins."""
    lib = build_mutant("none", None, None, str(tmp_path))
    env = dict(os.environ, MFC_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", PINS],
                       env=env, cwd=REPO, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
