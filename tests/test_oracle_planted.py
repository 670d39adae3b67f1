"""Each pin of the CPU oracle fails on a plausible mistake planted in the
oracle (oracle.planted: one deliberate error at a time, switched on in the
oracle itself, chmy_oracle.cpp enum Plant).  This is what makes the pins
pins: a check that still passed with the mistake in place would pin
nothing."""
import pytest

import numpy as np
import oracle as O
import test_oracle_closed_forms as CF
import test_oracle_masked as MK
import test_oracle_nonuniform as NU
import test_oracle_pins as PN
import test_oracle_reference_units as RU
import test_oracle_stokes_adjoint as SA

F64 = np.float64

# planted mistake -> (check, arguments) that must fail with it
CASES = [
    ("vertex_diff", CF.test_summation_by_parts_1d, ()),
    ("lerp_knots", CF.test_lerp_bilinear_exact_and_last_dim_outermost, ()),
    ("lerp_knots", RU.test_lerp_center_source, (F64,)),
    ("center_dirichlet", RU.test_bc_1d, (F64, O.CENTER)),
    ("neumann_sign", RU.test_bc_1d, (F64, O.CENTER)),
    ("bc_order", PN.test_bc_order_two_dirichlet_corner, ()),
    ("diff_flux", CF.test_diffusion_cosine_mode_decay, ()),
    ("stress_dev", PN.test_stress_update_affine_velocity_closed_form, (3, "scalar")),
    ("stress_dev", PN.test_stress_update_affine_velocity_closed_form, (2, "field")),
    ("pr_update", PN.test_stress_update_affine_velocity_closed_form, (3, "field")),
    ("edge_swap", PN.test_stress_viscosity_locations_affine_eta, (3,)),
    ("rv_swap", PN.test_velocity_residual_affine_stresses, (3,)),
    ("rv_swap", PN.test_velocity_residual_affine_stresses, (2,)),
    ("vel_ml_side", PN.test_velocity_coefficient_union_max, (2,)),
    ("vel_ml_side", PN.test_velocity_coefficient_union_max, (3,)),
    ("maxloc_2n", PN.test_maxloc_brute_force, (3,)),
    ("thermal_upwind", CF.test_thermal_pure_diffusion_and_upwind, ()),
    ("thermal_sign", CF.test_thermal_pure_diffusion_and_upwind, ()),
    ("sia_exp", PN.test_sia_flux_planar_bed_power_law, (3, (0.3, 0.0))),
    ("sia_grad", PN.test_sia_flux_planar_bed_power_law, (3, (0.12, 0.16))),
    ("sia_m", CF.test_sia_mass_balance, ()),
    ("resid_rp", CF.test_stokes_residual_matches_update_kernels, (3, "field")),
    ("grad_edge", SA.test_oracle_grad_formula_matches_central_differences, ((9, 7),)),
    ("pt_params", PN.test_pt_params_converge_at_the_damped_wave_rate, ()),
    ("vmag", RU.test_vmag_constant, (F64,)),
    ("divg", RU.test_divg_equals_sum_of_partials, (F64,)),
    ("mask_at_out", MK.test_masked_operators_are_unmasked_operators_of_the_masked_field, (2, (0, 1), 1, "fraction")),
    ("mask_at_out", MK.test_masked_sums_divg_lapl_divg_grad, (3, "fraction")),
    ("hlerp_arith", PN.test_hlerp_is_the_harmonic_rule, (1,)),
    ("hlerp_arith", PN.test_hlerp_is_the_harmonic_rule, (3,)),
    ("delta_sign", PN.test_delta_is_the_partial_before_its_inverse_spacing, (2,)),
    ("nu_diff_loc", NU.test_partial_scales_by_the_source_location_spacing, ()),
    ("nu_weight", NU.test_lerp_reproduces_affine_and_bilinear_fields, (1,)),
    ("nu_weight", NU.test_lerp_reproduces_affine_and_bilinear_fields, (2,)),
    ("nu_dirichlet", NU.test_boundary_ghosts_use_the_reference_indices, ()),
    ("nu_neumann", NU.test_boundary_ghosts_use_the_reference_indices, ()),
]


@pytest.mark.parametrize("plant,check,args", CASES, ids=[f"{p}-{c.__name__}" for p, c, _ in CASES])
def test_pin_fails_on_planted_mistake(plant, check, args):
    check(*args)                      # the restatement passes
    with O.planted(plant):
        with pytest.raises(AssertionError):
            check(*args)


def test_every_plant_is_exercised():
    assert {p for p, _, _ in CASES} == set(O.PLANTS)
