"""Configuration files in the reference's libconfig text
(c++/src/ConfigurationNetworkGrowth.cpp:210-296, ConfigurationNetworkTopology.cpp)."""
import pytest

from unsupervised_growth_amd import config as C
from unsupervised_growth_amd.topology import TopologyParams

TOPOLOGY_TEXT = ("spatialParameters = \n{\n\tarrangement = \"sphere\";\n\tA_RA2I = 5.500000;\n"
                 "\tSIGMA_RA2I = 0.125000;\n\tB_I2RA = 1.000000;\n\tSIGMA_I2RA = 30.000000;\n\n};\n\n")


def test_topology_text_and_round_trip(tmp_path):
    p = TopologyParams(arrangement="sphere", A_RA2I=5.5, SIGMA_RA2I=0.125, B_I2RA=1.0, SIGMA_I2RA=30.0)
    assert C.topology_configuration_text(p) == TOPOLOGY_TEXT
    C.write_topology_configuration(tmp_path / "network_parameters.cfg", p)
    q = C.read_topology_configuration(tmp_path / "network_parameters.cfg", C.ConnectionParams(0.3, 0.7, 2.0))
    assert (q.arrangement, q.A_RA2I, q.SIGMA_RA2I, q.B_I2RA, q.SIGMA_I2RA) == ("sphere", 5.5, 0.125, 1.0, 30.0)
    assert (q.Gei_max, q.Gie_max, q.delay_constant) == (0.3, 0.7, 2.0)


def test_growth_text_layout(tmp_path):
    c = C.GrowthConfig()
    txt = C.growth_configuration_text(c)
    lines = txt.split("\n")
    assert lines[:5] == ["synapticParameters = ", "{", "\tNss = 2;", "", "\tA_P = 2.000000;"]
    assert "\tTAU_D = 20.000000;\n\n\tBETA = 0.050000;\n\tBETA_SUPERSYNAPSE = 0.020000;\n\n" in txt
    assert "\tWEIGHT_MAX = 10.000000;\n\n};\n\nmaturationParameters = \n{\n\tE_GABA_MATURE = -80.000000;\n" in txt
    assert "\tRC_IMMATURE = 5.500000;\n\n\tDEATH_RATE_THRESHOLD = 0.010000;\n\n\tRATE_WINDOW_LONG = 2000;\n\n};" in txt
    assert "noiseParameters = \n{\n\twhite_noise_mean_soma = 0.000000;\n\twhite_noise_std_soma = 0.100000;\n" in txt
    assert txt.endswith("connectionParameters = \n{\n\tGei_max = 0.250000;\n\tGie_max = 0.000000;\n"
                        "\tdelay_constant = 0.000000;\n\n};\n\n")


def test_growth_round_trip_and_comments(tmp_path):
    c = C.GrowthConfig()
    c.synaptic.Nss, c.synaptic.T_0, c.maturation.RATE_WINDOW_LONG = 5, 1.25, 1500
    c.maturation.GCA_IMMATURE, c.noise.white_noise_std_dend, c.connection.delay_constant = 3.5, 0.2, 4.0
    C.write_growth_configuration(tmp_path / "g.cfg", c)
    d = C.read_growth_configuration(tmp_path / "g.cfg")
    assert d == c and isinstance(d.synaptic.Nss, int) and isinstance(d.maturation.RATE_WINDOW_LONG, int)
    text = ('// a comment\nsynapticParameters = { Nss = 3; A_P = 1e-1; # another\n WEIGHT_MAX = 7; };\n'
            '/* block\n comment */ noiseParameters : { white_noise_std_soma = .5; };\n')
    (tmp_path / "h.cfg").write_text(text)
    e = C.read_growth_configuration(tmp_path / "h.cfg")
    assert (e.synaptic.Nss, e.synaptic.A_P, e.synaptic.WEIGHT_MAX, e.noise.white_noise_std_soma) == (3, 0.1, 7.0, 0.5)
    with pytest.raises(ValueError):
        C.parse("synapticParameters = { Nss = ; };")
