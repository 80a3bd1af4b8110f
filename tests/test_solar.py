"""Solar geometry of the radiation hook: Spencer's series against known values."""
import math

import torch

from woof_nam_amd import solar


def test_declination_and_equation_of_time():
    # June solstice ~ +23.44 deg, December ~ -23.44, equinoxes ~ 0 (Spencer: within ~0.1 deg)
    for jd, want in (((2025, 6, 21, 0, 12, 0, 0, 0), 23.44), ((2025, 12, 21, 0, 12, 0, 0, 0), -23.44),
                     ((2025, 3, 20, 0, 12, 0, 0, 0), 0.0), ((2025, 9, 22, 0, 12, 0, 0, 0), 0.0)):
        gam, _ = solar.day_angle(jd)
        assert abs(math.degrees(solar.declination(gam)) - want) < 0.6, (jd, math.degrees(solar.declination(gam)))
    # equation of time: about -14 min mid-February, +16 min early November
    assert abs(solar.equation_of_time_minutes(solar.day_angle((2025, 2, 11, 0, 12, 0, 0, 0))[0]) + 14.2) < 1.0
    assert abs(solar.equation_of_time_minutes(solar.day_angle((2025, 11, 3, 0, 12, 0, 0, 0))[0]) - 16.4) < 1.0


def test_cosz_geometry():
    jd = (2025, 3, 20, 0, 0, 0, 0, 0)
    lat = torch.tensor([0.0, 0.0, math.radians(60.0), 0.0])
    lon = torch.tensor([0.0, math.radians(180.0), 0.0, math.radians(90.0)])
    cz = solar.cosz(jd, 12 * 3600.0, lat, lon)                      # 12 UTC at the equinox
    assert cz[0] > 0.99                                             # sun overhead at (0, 0)
    assert cz[1] < -0.99                                            # antipode: midnight
    assert abs(cz[2] - 0.5) < 0.02                                  # 60 N at noon: cos 60
    assert abs(cz[3]) < 0.06                                        # 90 E: sunset
    cm = solar.coszen(jd, 12 * 3600.0, 3600.0, lat, lon)
    assert torch.all(cm >= 0) and cm[0] > 0.98 and cm[1] == 0


def test_solar_constant_annual_cycle():
    jan = solar.solcon((2025, 1, 3, 0, 0, 0, 0, 0))
    jul = solar.solcon((2025, 7, 4, 0, 0, 0, 0, 0))
    assert 1.030 < jan / solar.S0 < 1.038 and 0.962 < jul / solar.S0 < 0.970
