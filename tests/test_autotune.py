"""The autotuner's table round-trips through the library (CPU: no compute)."""
import json

import numpy as np


def test_apply_and_load_table(tmp_path):
    from matrixinversion_amd import api, autotune

    layout, grab = api.tuning_get()
    try:
        t = {"layout": [0] * 33, "grab": [3] * 33}
        t["layout"][5] = 1
        p = tmp_path / "t.json"
        p.write_text(json.dumps(t))
        autotune.load_table(str(p))
        l2, g2 = api.tuning_get()
        assert l2[5] == 1 and l2[6] == 0 and (g2 == 3).all()
    finally:
        api.tuning_set(layout, grab)
    assert (api.tuning_get()[0] == layout).all()


def test_default_table():
    from matrixinversion_amd import api

    layout, grab = api.tuning_get()
    assert (layout[1:13] == 1).all() and (layout[13:] == 0).all()
    assert (grab[1:13] == 1).all() and (grab[13:32] == 8).all() and grab[32] == 4
    assert np.all(grab >= 1)
