"""The oracle's pins catch plausible mistakes.  CPU only.

Each case plants one mistake into a copy of oracle/gmc_oracle.c in a temporary
directory (a dropped min() branch, the Mach row-zeroing removed, a flipped
H.J sign, a wrong Giordano coefficient slot, ...), builds that copy with gcc,
and runs the named pins of test_oracle_pins.py or test_oracle_magma.py
against it in a fresh process (GMC_ORACLE_LIB points
orc.py at the copy).  The pins must fail.  The oracle itself is untouched.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (id, original text, planted mistake, pins that must catch it)
MUTANTS = [
    ("tau_viscous_branch_dropped",
     "td[1] = dmin(use, h * h * rho / (12.0 * q->pr.mu));", "td[1] = use;",
     "test_tau_2019_branches"),
    ("tau_thermal_min_is_max",
     "td[iE] = dmin(use, h * h * rho * q->pr.cv / (12.0 * q->pr.kappa));",
     "td[iE] = fmax(use, h * h * rho * q->pr.cv / (12.0 * q->pr.kappa));",
     "test_tau_2019_branches"),
    ("tau_steady_ignored",
     "if (steady) use = h / (2.0 * (v_nrm + c)) + h / (2.0 * v_nrm);",
     "if (0) use = 0.0;",
     "test_tau_2019_branches"),
    ("mach_row_zeroing_dropped",
     "for (int j = 1; j < n; j++) tau[0][j] = 0.0;", "(void)0;",
     "test_tau_2019_low_mach_correction"),
    ("mach_threshold_wrong",
     "if (v_nrm / c < 0.1) {", "if (v_nrm / c < 1.0) {",
     "test_tau_2019_branches"),
    ("mach_trace_2d_only_in_3d",
     "for (int k = 0; k < dim; k++) trG += g->G[k][k];\n    tau[0][0] = pow(pow(",
     "for (int k = 0; k < 2; k++) trG += g->G[k][k];\n    tau[0][0] = pow(pow(",
     "test_3d_point_reduces_to_2d"),
    ("compute_h_acceleration_dropped",
     "for (int i = 0; i < dim; i++) al[1] += s1[i] * Gs[i];", "(void)0;",
     "test_compute_h_closed_forms_2d or test_compute_h_through_the_element"),
    ("compute_h_binormal_dropped",
     "for (int i = 0; i < 3; i++) al[2] += s2[i] * Gs[i];", "(void)0;",
     "test_compute_h_closed_forms_3d"),
    ("dc_length_without_abs",
     "h += fabs(s);", "h += s;",
     "test_dc_nu_on_a_state_varying_along_x"),
    ("dc_sharp1_power",
     "if (a == 1.0) nu = zn * h * pow(use1, -0.5);", "if (a == 1.0) nu = zn * h * pow(use1, -1.0);",
     "test_dc_nu_on_a_state_varying_along_x"),
    ("boundary_HJ_sign_flipped",
     "q->Fdif[k][iE] = e - qf[k] - HJ;", "q->Fdif[k][iE] = e - qf[k] + HJ;",
     "test_boundary_residual_one_edge_kat"),
    ("neumann_q_ignored",
     "if (nb->q_mask & (1u << k)) qf[k] = nb->q[k];", "(void)0;",
     "test_neumann_wall_overrides"),
    ("side_weight_wrong",
     "double orc_w_bd(int dim) { return dim == 2 ? 1.0 / 2.0 : 1.0 / 6.0; }",
     "double orc_w_bd(int dim) { return dim == 2 ? 1.0 : 1.0 / 6.0; }",
     "test_global_conservation_on_a_mesh"),
    ("tau2007_g_dot_g_wrong",
     "for (int k = 0; k < dim; k++) gg += gvec[k] * gvec[k];",
     "for (int k = 0; k < dim; k++) gg += gvec[k] * gvec[k] * 2.0;",
     "test_tau_2007_closed_form"),
    ("tau2014_row0_correction_dropped",
     "td[0] = pow(td[0] + rho * td[1] * trG, -1);", "(void)0;",
     "test_tau_2014_closed_form"),
    ("tau2001_uses_2019_viscous_scale",
     "td[1] = dmin(use, rho * h * h / (q->pr.mu * 12));", "td[1] = dmin(use, rho * h * h / (q->pr.mu * 6));",
     "test_tau_2001_unsteady"),
    ("mass_flux_energy_term",
     "q->Fadv[k][iE] = m * E + p * v[k];", "q->Fadv[k][iE] = m * E;",
     "test_mass_flux_side_kat"),
    ("mass_flux_jacobian_entry",
     "A[iE][iT] = m * de_dT;", "A[iE][iT] = m * pr->cp;",
     "test_mass_flux_side_matrix_is_jacobian"),
    ("interior_HJ_sign_flipped",
     "Kk[iE][iY + i] = rho * pr->D_comp[i] * H[i];", "Kk[iE][iY + i] = -rho * pr->D_comp[i] * H[i];",
     "test_interior_diffusive_flux_closed_form"),
]

# mistakes in the magma submixture, caught by test_oracle_magma.py
MAGMA_MUTANTS = [
    ("magma_interp_from_first_node",
     "*out = y[i] + slope * (p - pt[i]);", "*out = y[i] + slope * (p - pt[0]);",
     "test_lookup_linear"),
    ("magma_bilinear_wrong_corner",
     "(T - T1) * (p2 - p) * m[(j + 1) * np + i]", "(T - T1) * (p2 - p) * m[j * np + i]",
     "test_lookup_bilinear"),
    ("magma_last_node_accepted",
     "if (!(f >= 0.0) || f + 1 >= (double)n) return GMC_ERR_RANGE;",
     "if (!(f >= 0.0) || f >= (double)n) return GMC_ERR_RANGE;",
     "test_lookup_linear"),
    ("magma_gas_phase_from_total",
     "wf[3] = wf_comp[1] - wf[1];", "wf[3] = wf_comp[1];",
     "test_phase_split"),
    ("melt_beta_sign",
     "o->beta = -1. / mv * m->dV_dp;", "o->beta = 1. / mv * m->dV_dp;",
     "test_chemicals_thermodynamics"),
    ("dissolved_alpha_term",
     "o->alpha = a[1] + 2 * a[2] * T + 3 * a[3] * T * T", "o->alpha = a[1] + a[2] * T + 3 * a[3] * T * T",
     "test_chemicals_thermodynamics"),
    ("melt_cp_by_mass",
     "for (int i = 0; i < 10; i++) m->cp += x[i] * cpo[i];",
     "for (int i = 0; i < 10; i++) m->cp += m->wf[i] * cpo[i];",
     "test_melt_lange_carmichael"),
    ("crystal_density_dropped",
     "rho = (1.0 - g->crystal_vf) * rho + g->crystal_vf * g->crystal_rho;", "(void)0;",
     "test_crystal_density_blend"),
    ("wood_sound_speed_power",
     "snd += wf[c] / (pow(ch[c].rho, 2) * pow(ch[c].c, 2));", "snd += wf[c] / (pow(ch[c].rho, 2) * ch[c].c);",
     "test_mixture_rules_and_wood"),
    ("giordano_fe_not_zeroed",
     "  w[3] = 0.0;\n", "  (void)0;\n",
     "test_giordano_fe_total"),
    ("giordano_water_b6_dropped",
     "(b[5] + b[6]) * h2o_pct", "b[6] * h2o_pct",
     "test_giordano_water"),
    ("giordano_c_slot_swapped",
     "c[2] * (mp[4] + mp[5] + mp[6]) + c[3] * mp[7]", "c[3] * (mp[4] + mp[5] + mp[6]) + c[2] * mp[7]",
     "test_giordano_anhydrous"),
    ("ishii_zuber_sign",
     "mu = mu / (1.0 - o->vf_gas);", "mu = mu / (1.0 + o->vf_gas);",
     "test_gas_and_crystal_corrections"),
    ("pal_half_newton_step",
     "    x -= F / Fp;\n", "    x -= 0.5 * F / Fp;\n",
     "test_pal_2003"),
    ("interior_shear_rate_dropped",
     "    orc_properties(P, Q.p, Q.T, Q.Y, sr, &Q.pr);\n    if (Q.pr.st) return Q.pr.st;\n    flux_interior",
     "    orc_properties(P, Q.p, Q.T, Q.Y, 0.0, &Q.pr);\n    if (Q.pr.st) return Q.pr.st;\n    flux_interior",
     "test_interior_point_viscosity_at_its_shear_rate"),
    ("magma_D_not_the_submixture",
     "    *D = m.D;\n    return 0;", "    *D = c->diffusivity * 2.0;\n    return 0;",
     "test_magma_component_in_the_mixture"),
]


@pytest.fixture(scope="module")
def source():
    with open(os.path.join(ORACLE, "gmc_oracle.c")) as f:
        return f.read()


CASES = [m + ("test_oracle_pins.py",) for m in MUTANTS] + [m + ("test_oracle_magma.py",) for m in MAGMA_MUTANTS]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("mid,old,new,pins,tests", CASES, ids=[m[0] for m in CASES])
def test_pins_catch_planted_mistake(mid, old, new, pins, tests, source, tmp_path):
    assert source.count(old) == 1, f"mutation site of {mid} not found exactly once"
    src = tmp_path / "gmc_oracle.c"
    src.write_text(source.replace(old, new))
    lib = tmp_path / "liboracle_mut.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared", "-I", ORACLE,
                    "-o", str(lib), str(src), "-lm"], check=True)
    env = dict(os.environ, GMC_ORACLE_LIB=str(lib))
    res = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", tests),
                          "-q", "-x", "-p", "no:cacheprovider", "-k", pins],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    out = res.stdout + res.stderr
    # exit code 1: tests ran and some failed (2+ would be a collection or usage error)
    assert res.returncode == 1 and " failed" in out, out[-2000:]
