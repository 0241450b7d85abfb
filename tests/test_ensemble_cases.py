"""The seeded ensemble of 2D cylinder cases: sizes from the stated sets, Re in [40, 300], the
same cases for the same seed, and per-case grids with the single 2D case's stretching law."""
import numpy as np

from fvm_amd.grid import (ENSEMBLE_NI, ENSEMBLE_NJ, ensemble_cases, stretch_ratio)


def test_ensemble_cases_seeded_and_in_range():
    a, b = ensemble_cases(12), ensemble_cases(12)
    sizes = set()
    for ca, cb in zip(a, b):
        assert (ca["ni"], ca["nj"], ca["re"]) == (cb["ni"], cb["nj"], cb["re"])
        assert np.array_equal(ca["ic"]["u"], cb["ic"]["u"])
        assert ca["ni"] in ENSEMBLE_NI and ca["nj"] in ENSEMBLE_NJ and 40.0 <= ca["re"] <= 300.0
        assert ca["dt"] == 0.005 * min(1.0, 256.0 / (ca["nj"] - 1))
        assert ca["xg"].shape == (ca["ni"], ca["nj"], 1)
        sizes.add((ca["ni"], ca["nj"]))
    assert len(sizes) > 4  # uneven sizes
    # the first radial spacing is the wall spacing h1 and the last node sits on r_out
    c = a[0]
    r = np.hypot(c["xg"][:, 0, 0], c["yg"][:, 0, 0])
    assert abs(r[0] - 0.5) < 1e-12 and abs(r[1] - r[0] - 0.002) < 1e-12
    assert abs(r[-1] - 30.0) < 1e-9
    q = stretch_ratio(c["ni"] - 1, 0.002, 0.5, 30.0)
    assert abs((r[2] - r[1]) / (r[1] - r[0]) - q) < 1e-9
    # distinct cases draw distinct noise
    assert not np.array_equal(a[0]["ic"]["u"][1:3, 1:3], a[1]["ic"]["u"][1:3, 1:3])
