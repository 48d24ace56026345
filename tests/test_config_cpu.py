"""wflowhip.config: the reference's TOML [time] / [model] settings mapped onto wfh_config
(Wflow/src/config_structure.jl:28-110), on the [time] and [model] sections of the reference's
own test configuration (tests/golden/sbm_config_model.toml, from Wflow/test/sbm_config.toml)."""
import os

import pytest

from wflowhip import config

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reference_test_configuration():
    c, info = config.load(os.path.join(HERE, "golden", "sbm_config_model.toml"))
    assert c.dt == 86400.0
    assert c.kinematic_wave_adaptive == 1 and c.snow == 1 and c.snow_gravitational_transport == 1
    assert c.n_config_layers == 3 and list(c.soil_layer_thickness_mm[:3]) == [100.0, 300.0, 800.0]
    assert (c.land_dt, c.river_dt, c.subsurface_dt) == (3600.0, 900.0, 86400.0)  # defaults
    assert c.subsurface_alpha_coefficient == 1.0 and c.water_mass_balance == 0
    assert info["reservoirs"] and info["cold_start"] and not info["exponential_constant"]


def test_defaults_and_profiles():
    c, info = config.load(text='[time]\ntimestepsecs = 3600\n[model]\ntype = "sbm"\n'
                               'saturated_hydraulic_conductivity_profile = "exponential_constant"\n'
                               'snow_gravitational_transport__flag = true\n')
    assert c.dt == 3600.0 and c.snow == 0 and c.snow_gravitational_transport == 0  # no snow model: no-op
    assert info["exponential_constant"] and not info["reservoirs"]


def test_floodplain_flag():
    c, info = config.load(text='[time]\ntimestepsecs = 86400\n[model]\ntype = "sbm"\nfloodplain_1d__flag = true\n')
    assert c.floodplain_1d == 1 and info["floodplain"]
    c, info = config.load(text='[time]\ntimestepsecs = 86400\n[model]\ntype = "sbm"\n')
    assert c.floodplain_1d == 0 and not info["floodplain"]


@pytest.mark.parametrize("model", ['type = "sediment"', 'type = "sbm"\nglacier__flag = true',
                                   'type = "sbm"\nriver_routing = "local_inertial"',
                                   'type = "sbm"\nsaturated_hydraulic_conductivity_profile = "layered"',
                                   'type = "sbm"\nsome_unknown__flag = true'])
def test_unsupported_settings_are_rejected(model):
    with pytest.raises(ValueError):
        config.load(text="[time]\ntimestepsecs = 86400\n[model]\n" + model + "\n")
