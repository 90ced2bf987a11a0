"""Mutation checks of the oracle's pins (CPU only; the oracle is the checker).

One mistake at a time is planted in a copy of the oracle's C++ source in a
temporary directory, the copy is built with the oracle's own Makefile, and
tests/test_oracle.py runs against it in a subprocess.  Each mistake must fail
the tests listed as catching it, and no others; where a mistake keeps the
scheme consistent (conservative, free-stream preserving, mirror symmetric),
the whole-step property tests must still pass — which is why the value pins
that catch it exist.  The planted text must occur exactly once in the
source, so an edit of the oracle that moves it fails here until this table
follows.
"""
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# whole-step properties of the reference step that a reconstruction mistake
# made alike on both sides of every face does not break
PROPERTIES = ["test_free_stream_preserved_periodic_and_boundaries", "test_conservation_periodic_inviscid_and_viscous",
              "test_steps_commute_with_the_diagonal_mirror"]

# name: (source file, [(text, planted text)], tests that fail, tests that still pass)
MUTATIONS = {
    # computeslope2d.jl:118-126: UpLeft = uL + 0.5 lim(...), UpRight = uR - 0.5 lim(...)
    "muscl_half_to_one": (
        "oracle.cpp",
        [("UpLeft[v] = uLeftp[v] + 0.5 * limiter", "UpLeft[v] = uLeftp[v] + 1.0 * limiter"),
         ("UpRight[v] = uRightp[v] - 0.5 * limiter", "UpRight[v] = uRightp[v] - 1.0 * limiter")],
        ["test_reference_muscl_linear_field_known_answer", "test_reference_residual_order_of_accuracy"], PROPERTIES),
    "muscl_up_down_swapped": (
        "oracle.cpp",
        [("limiter(prm->limiter, uLeftp[v] - uDownp[v]", "limiter(prm->limiter, uLeftp[v] - uUpp[v]"),
         ("uUpp[v] - uRightp[v], ksi)", "uDownp[v] - uRightp[v], ksi)")],
        ["test_reference_muscl_linear_field_known_answer", "test_reference_residual_order_of_accuracy"], PROPERTIES),
    "limiter_ksi_1e-6": (
        "oracle.cpp", [("const double ksi = prm->limiter_eps;", "const double ksi = 1e-6;")],
        ["test_reference_muscl_linear_field_known_answer"], PROPERTIES),
    # thermo.jl:17-40, c_p polynomials
    "cp_low_T4_exponent": (
        "oracle.cpp", [("6.272365e-10 * lit_pow_4(T)", "6.272365e-11 * lit_pow_4(T)")],
        ["test_air_conductivity_known_answers_per_branch"], ["test_air_conductivity_structure_and_branches"]),
    "cp_high_T3_digit": (
        "oracle.cpp", [("-3.075525e-9 * (T * T * T)", "-3.075525e-8 * (T * T * T)")],
        ["test_air_conductivity_known_answers_per_branch", "test_air_conductivity_structure_and_branches"], []),
    # AUSMflux2dFast.jl: beta = 1/8 in M_beta
    "ausm_beta_1_16": (
        "oracle.cpp",
        [("const double m_dot12 = M_p((UL * nx + VL * ny) / a12, 1.0 / 8.0) + M_m((UR * nx + VR * ny) / a12, 1.0 / 8.0);",
          "const double m_dot12 = M_p((UL * nx + VL * ny) / a12, 1.0 / 16.0) + M_m((UR * nx + VR * ny) / a12, 1.0 / 16.0);")],
        ["test_ausm_plus_subsonic_known_answer", "test_reference_gpu_stage_fp64"], PROPERTIES),
    # utilsFVM2dp.jl:305-320, weighted node means
    "node_means_unweighted": (
        "oracle.cpp", [("const double wi = D(m->node_stencils, nn, J, j);", "const double wi = 1.0;")],
        ["test_node_interpolation_uses_the_stencil_weights"], ["test_node_interpolation_constant_and_convex"]),
    # calcDiffTerm.jl: the energy row carries k
    "diff_energy_without_k": (
        "oracle.cpp", [("Udiff[3 * n + i - 1] = mu * k * le;", "Udiff[3 * n + i - 1] = mu * le;")],
        ["test_diff_term_scales_with_mu_and_k"], ["test_diff_term_laplacian_converges"]),
    # bcInviscidWall.jl: the slip reflection's n_x n_y terms
    "slip_nxny_sign": (
        "oracle.cpp",
        [("bnd[1] = cur[1] * (ny * ny - nx * nx) - 2.0 * nx * ny * cur[2];",
          "bnd[1] = cur[1] * (ny * ny - nx * nx) + 2.0 * nx * ny * cur[2];"),
         ("bnd[2] = cur[2] * (nx * nx - ny * ny) - 2.0 * nx * ny * cur[1];",
          "bnd[2] = cur[2] * (nx * nx - ny * ny) + 2.0 * nx * ny * cur[1];")],
        ["test_ref_slip_isothermal_fixed_and_zero_gradient_ghosts", "test_slip_annulus_conserves_mass_and_energy"],
        ["test_slip_walls_conserve_mass_and_energy"]),
    "isothermal_ghost_T_is_Tw": (
        "oracle.cpp", [("double Tg = 2.0 * b->data[0] - Tc;", "double Tg = b->data[0];")],
        ["test_ref_slip_isothermal_fixed_and_zero_gradient_ghosts"], []),
    # the ns scheme (not in the reference): kappa = mu c_p / Pr, the 2/3 of the
    # stress, the LSQ weights, the viscous term of the time step
    "ns_kappa_mu_cp_times_Pr": (
        "oracle_ns.cpp", [("const R kap = mu * cp / (R)p->Pr;", "const R kap = mu * cp * (R)p->Pr;")],
        ["test_ns_isothermal_wall_heat_flux_known_answer", "test_ns_viscous_stress_and_heat_flux_known_answers"],
        ["test_steps_commute_with_the_diagonal_mirror"]),
    "ns_two_thirds_in_txx_only": (
        "oracle_ns.cpp",
        [("const R txx = mu * (2 * gf[0][0] - (R)(2.0 / 3.0) * div);",
          "const R txx = mu * (2 * gf[0][0] - (R)(1.0 / 3.0) * div);")],
        ["test_ns_viscous_stress_and_heat_flux_known_answers", "test_steps_commute_with_the_diagonal_mirror"], []),
    "ns_lsq_wy_for_wx": (
        "oracle_ns.cpp", [("gx[v] += (R)c.wx[k] * d;", "gx[v] += (R)c.wy[k] * d;")],
        ["test_ns_lsq_exact_for_linear_fields", "test_ns_venkatakrishnan_known_answer",
         "test_ns_viscous_stress_and_heat_flux_known_answers", "test_steps_commute_with_the_diagonal_mirror"], []),
    "ns_dt_viscous_term_halved": (
        "oracle_ns.cpp",
        [("lam += 2.0 * p->gamma * sutherland<double>(T, p) / (p->Pr * q[0] * G[i].hmin);",
          "lam += 1.0 * p->gamma * sutherland<double>(T, p) / (p->Pr * q[0] * G[i].hmin);")],
        ["test_ns_time_step_known_answer"], ["test_ns_free_stream_and_conservation_periodic"]),
}


def _run_mutant(name, tmp):
    """Build the oracle with mistake `name` planted and run tests/test_oracle.py
    against it; returns ({failed tests}, {passed tests}, summary line)."""
    fname, reps, _, _ = MUTATIONS[name]
    d = os.path.join(tmp, name)
    shutil.copytree(os.path.join(ROOT, "oracle"), os.path.join(d, "oracle"),
                    ignore=shutil.ignore_patterns("*.so", "__pycache__", "_ref"))
    path = os.path.join(d, "oracle", fname)
    src = open(path).read()
    for a, b in reps:
        assert src.count(a) == 1, f"{name}: {a!r} occurs {src.count(a)} times in oracle/{fname}"
        src = src.replace(a, b)
    open(path, "w").write(src)
    subprocess.run(["make", "-s", "-C", os.path.join(d, "oracle"), "CXXFLAGS=-O1 -fPIC -fopenmp -ffp-contract=off "
                    "-fno-fast-math -std=c++17 -w"], check=True, capture_output=True)
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([d, ROOT]), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-rA", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_oracle.py")], cwd=tmp, env=env, capture_output=True,
                       text=True, timeout=600)
    failed = set(re.findall(r"^FAILED \S+::(\w+)", r.stdout, re.M))
    passed = set(re.findall(r"^PASSED \S+::(\w+)", r.stdout, re.M))
    lines = r.stdout.strip().splitlines()
    return failed, passed, lines[-1] if lines else r.stderr[-500:]


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        futs = {n: ex.submit(_run_mutant, n, tmp) for n in MUTATIONS}
        return {n: f.result() for n, f in futs.items()}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_each_planted_mistake_is_caught_by_its_pins(mutants, name):
    failed, passed, summary = mutants[name]
    _, _, catchers, still_pass = MUTATIONS[name]
    print(name, summary, sorted(failed))
    assert failed == set(catchers), (name, sorted(failed))
    assert set(still_pass) <= passed, (name, sorted(set(still_pass) - passed))
