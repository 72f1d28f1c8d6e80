"""The pins of tests/test_oracle_pins.py fail on plausible mistakes.

Each case plants one mistake in a copy of the oracle's source (the C
restatement oracle/iisph_oracle.c, or the learned-start restatement
oracle/learn_oracle.py), builds it on the CPU, and runs the pins that cover
that function against the copy: at least one of them must fail. The
unmodified source passes every pin (tests/test_oracle_pins.py), so a pin that
still passed here would be checking nothing the mistake changes.
"""
from __future__ import annotations

import subprocess
import types
from pathlib import Path

import pytest

import test_hydrostatic as hydro
import test_neighbor_properties as nprops
import test_oracle_pins as pins

ORACLE = Path(__file__).resolve().parents[1] / "oracle"

# (name, original text, planted text, pins that must catch it). Every
# original text occurs exactly once in its file.
C_MUTANTS = [
    ("rho_adv: v_adv,i + v_b (sign of the boundary velocity)",
     "{va[0] - vbb[0], va[1] - vbb[1], va[2] - vbb[2]}",
     "{va[0] + vbb[0], va[1] + vbb[1], va[2] + vbb[2]}",
     ["pin_rho_adv_boundary_pair"]),
    ("rho_adv: v in place of v_adv",
     "*va = v_adv + 3 * (size_t)i;", "*va = v + 3 * (size_t)i;",
     ["pin_rho_adv_boundary_pair"]),
    ("rho_adv: psi_b dropped",
     "t2 += dot3(vib, gw) * psi[b];", "t2 += dot3(vib, gw);",
     ["pin_rho_adv_boundary_pair"]),
    ("rho_adv: v_adv,j - v_adv,i (sign of the fluid term)",
     "{va[0] - vk[0], va[1] - vk[1], va[2] - vk[2]}",
     "{vk[0] - va[0], vk[1] - va[1], vk[2] - va[2]}",
     ["pin_rho_adv_linear_field"]),
    ("XSPH: v_i - v_j",
     "acc[c] + (vk[c] - vi[c]) * coef", "acc[c] + (vi[c] - vk[c]) * coef",
     ["pin_xsph_three_particles", "pin_xsph_lattice_fields"]),
    ("XSPH: m / rho_i in place of m / rho_j",
     "const float coef = m / rho[k] * spline_W", "const float coef = m / rho[i] * spline_W",
     ["pin_xsph_three_particles"]),
    ("Laplacian: sign of the fluid term",
     "t1[c] = t1[c] - gw[c] * dd / rho[k] / r2", "t1[c] = t1[c] + gw[c] * dd / rho[k] / r2",
     ["pin_laplacian_three_particles"]),
    ("Laplacian: rho_i in place of rho_j",
     "t1[c] = t1[c] - gw[c] * dd / rho[k] / r2", "t1[c] = t1[c] - gw[c] * dd / rho[i] / r2",
     ["pin_laplacian_three_particles"]),
    ("Laplacian: psi_b dropped",
     "t2[c] = t2[c] - gw[c] * psi[b] * dd / r2", "t2[c] = t2[c] - gw[c] * dd / r2",
     ["pin_laplacian_boundary_pair"]),
    ("update_p: sign(0) = -1",
     "a = (a < 0.0f ? -1.0f : 1.0f) * inv_small", "a = (a <= 0.0f ? -1.0f : 1.0f) * inv_small",
     ["pin_update_p_isolated"]),
    ("update_p: guard band ignores the sign of a_ii",
     "a = (a < 0.0f ? -1.0f : 1.0f) * inv_small", "a = inv_small",
     ["pin_update_p_guard_and_clamp"]),
    ("update_p: omega dropped",
     "(rho0 - rho_adv[i] - Ap[i]) * omega / a", "(rho0 - rho_adv[i] - Ap[i]) / a",
     ["pin_update_p_isolated"]),
    ("update_p: no clamp at zero",
     "p[i] = pi > 0.0f ? pi : 0.0f;", "p[i] = pi;",
     ["pin_update_p_guard_and_clamp", "pin_jacobi_fixed_point"]),
    ("update_p: rho_adv + A p in place of rho0 - rho_adv - A p",
     "(rho0 - rho_adv[i] - Ap[i]) * omega / a", "(rho0 - rho_adv[i] + Ap[i]) * omega / a",
     ["pin_jacobi_fixed_point"]),
]

# mistakes of the neighbour search (IISPH.py:52-163), caught by brute force
NEIGHBOR_MUTANTS = [
    ("distance test r <= h (distances one ulp short)",
     "return sqrtf(dx * dx + dy * dy + dz * dz);",
     "return nextafterf(sqrtf(dx * dx + dy * dy + dz * dz), 0.0f);",
     ["test_fluid_sets_equal_brute_force", "test_boundary_sets_equal_brute_force"]),
    ("distance summed as dx^2 + (dy^2 + dz^2)",
     "return sqrtf(dx * dx + dy * dy + dz * dz);",
     "return sqrtf(dx * dx + (dy * dy + dz * dz));",
     ["test_fluid_sets_equal_brute_force", "test_boundary_sets_equal_brute_force"]),
    ("query range not clamped to dims cells",
     "q->xe = xe < q->xs + g->dx - 1 ? xe : q->xs + g->dx - 1;", "q->xe = xe;",
     ["test_fluid_sets_equal_brute_force", "test_boundary_sets_equal_brute_force"]),
    ("query range clamped one cell short on y",
     "q->ye = ye < q->ys + g->dy - 1 ? ye : q->ys + g->dy - 1;",
     "q->ye = ye < q->ys + g->dy - 1 ? ye : q->ys + g->dy - 2;",
     ["test_fluid_sets_equal_brute_force", "test_boundary_sets_equal_brute_force"]),
]

# mistakes of the whole step that only physics catches: the settled layer's
# pressure gradient must balance gravity (tests/test_hydrostatic.py)
PHYSICS_MUTANTS = [
    ("kick: sum_j d_ij p_j dropped from the pressure velocity",
     "(d_ii[3 * (size_t)i + c] * p[i] + s[3 * (size_t)i + c])",
     "(d_ii[3 * (size_t)i + c] * p[i])",
     ["test_oracle_settled_layer_is_hydrostatic"]),
]

PY_MUTANTS = [
    ("features: the particle's own term left out of the means",
     "sw = w0 + w.sum()\n        f[i, 4:] = (w0 * f[i, :4] + w @ f[j, :4]) / sw",
     "sw = w.sum()\n        f[i, 4:] = (w @ f[j, :4]) / sw",
     ["pin_learned_means"]),
    ("features: means weighted by W^2",
     "f[i, 4:] = (w0 * f[i, :4] + w @ f[j, :4]) / sw",
     "f[i, 4:] = (w0 * w0 * f[i, :4] + (w * w) @ f[j, :4]) / (w0 * w0 + (w * w).sum())",
     ["pin_learned_means"]),
    ("arbiter: weight not capped at 1",
     "w = np.minimum(np.abs(np.asarray(a_ii, np.float64) / a_ref(cfg, dt)), 1.0)",
     "w = np.abs(np.asarray(a_ii, np.float64) / a_ref(cfg, dt))",
     ["pin_learned_arbiter"]),
    ("a_ref: 1/h dropped from |W'| of the outer piece",
     "sigma * 6 * (1 - q) ** 2 / h)", "sigma * 6 * (1 - q) ** 2)",
     ["pin_learned_lattice_constant"]),
]


def _mutate(text: str, old: str, new: str) -> str:
    assert text.count(old) == 1, f"mutation site not unique: {old!r}"
    return text.replace(old, new)


def _expect_caught(run_pins, names, module=pins):
    caught = []
    for name in names:
        try:
            run_pins(getattr(module, name))
        except AssertionError:
            caught.append(name)
    assert caught, f"no pin among {names} caught the planted mistake"


def _build_mutant(tmp_path, old, new):
    src = _mutate((ORACLE / "iisph_oracle.c").read_text(), old, new)
    c = tmp_path / "iisph_oracle.c"
    c.write_text(src.replace('#include "iisph_oracle.h"', f'#include "{ORACLE}/iisph_oracle.h"')
                 .replace('#include "../include/wsph.h"',
                          f'#include "{ORACLE.parent}/include/wsph.h"'))
    so = tmp_path / "liborc_mutant.so"
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-shared", "-o", str(so), str(c), str(ORACLE / "mesh_oracle.c"),
                    "-lm"], check=True, capture_output=True)
    return so


@pytest.mark.parametrize("name,old,new,names", C_MUTANTS, ids=[m[0] for m in C_MUTANTS])
def test_c_oracle_pins_catch(orc, tmp_path, name, old, new, names):
    so = _build_mutant(tmp_path, old, new)
    with orc.using(so):
        _expect_caught(lambda pin: pin(orc), names)


@pytest.mark.parametrize("name,old,new,names", NEIGHBOR_MUTANTS,
                         ids=[m[0] for m in NEIGHBOR_MUTANTS])
def test_neighbor_properties_catch(orc, tmp_path, name, old, new, names):
    so = _build_mutant(tmp_path, old, new)
    with orc.using(so):
        _expect_caught(lambda test: test(orc), names, nprops)


@pytest.mark.parametrize("name,old,new,names", PHYSICS_MUTANTS,
                         ids=[m[0] for m in PHYSICS_MUTANTS])
def test_hydrostatic_catches(orc, tmp_path, name, old, new, names):
    so = _build_mutant(tmp_path, old, new)
    with orc.using(so):
        _expect_caught(lambda test: test(orc), names, hydro)


@pytest.mark.parametrize("name,old,new,names", PY_MUTANTS, ids=[m[0] for m in PY_MUTANTS])
def test_learn_oracle_pins_catch(name, old, new, names):
    src = _mutate((ORACLE / "learn_oracle.py").read_text(), old, new)
    mod = types.ModuleType("learn_oracle_mutant")
    exec(compile(src, "learn_oracle_mutant.py", "exec"), mod.__dict__)
    _expect_caught(lambda pin: pin(mod), names)
