"""Each oracle pin catches a plausible mistake: with one known mistake planted in
the oracle (pyoracle.plant, the oracle's PLANT_* codes), the pin written for
that function fails, and with none planted it passes (the rest of the suite).

The mistakes are the slips the pins were chosen against: a wrong multiple
importance weight, a dropped cosine, a wrong cone solid angle or a cone sampled
non-uniformly in cos(theta), a bounce weighted by the wrong pdf, Russian
roulette without its reweighting, an off-by-one or transposed texel lookup,
Kahan accumulation without its compensation, an untransposed normal matrix, a
background lookup with its atan2 arguments swapped, adaptive sampling's divisor
off by one, an unnormalised normal-mapped normal and ambient light without the
albedo.  The FP32 sin/cos/atan2 routine is shared with the kernels and is not
planted in; its pins compare it with FP64 directly, ulp by ulp.
"""
import pytest

import pyoracle as O
import test_oracle_features as F
import test_oracle_pins as P

CASES = [
    ("mis_balance", P.test_sphere_light_plate_closed_form, ()),
    ("drop_cosine", P.test_sphere_light_plate_closed_form, ()),
    ("cone_omega", P.test_sphere_light_cone_sampling_and_pdf, ()),
    ("cone_omega", P.test_sphere_light_plate_closed_form, ()),
    ("cone_square", P.test_sphere_light_cone_sampling_and_pdf, ()),
    ("bounce_pdf", P.test_one_bounce_floor_wall_quadrature, (False,)),
    ("rr_no_reweight", P.test_one_bounce_floor_wall_quadrature, (True,)),
    ("tex_off_by_one", P.test_bilinear_texture_against_scipy_and_hand_values, ()),
    ("tex_transposed", P.test_bilinear_texture_against_scipy_and_hand_values, ()),
    ("kahan_off", P.test_kahan_sum_against_exact, ()),
    ("normal_matrix", P.test_instanced_fp64_equals_pretransformed_fp64, ()),
    ("background_swap", F.test_background_on_miss, (64,)),
    ("adaptive_divisor", F.test_adaptive_sampling_stops_and_divisor, ()),
    ("nmap_unnormalised", F.test_normal_map_tilts_the_shading_normal, (64,)),
    ("ambient_no_albedo", F.test_ambient_only, (64,)),
]


@pytest.mark.parametrize("mistake,pin,args", CASES, ids=[f"{m}-{p.__name__}" for m, p, _ in CASES])
def test_planted_mistake_fails_its_pin(mistake, pin, args):
    O.plant(mistake)
    try:
        with pytest.raises(AssertionError):
            pin(*args)
    finally:
        O.plant(None)


def test_every_planted_mistake_has_a_case():
    assert {m for m, _, _ in CASES} == set(O.PLANTS)
