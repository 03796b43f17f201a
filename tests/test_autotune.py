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
    """The built-in table: the one the mixed-batch tuner confirmed (quads up
    to n = 12, 8 row tiles per ragged work item, 16 at n = 30, 4 at n = 32),
    LDS rows padded for n % 4 == 0 except n = 28."""
    from matrixinversion_amd import api

    layout, grab = api.tuning_get()
    assert (layout[1:13] == 1).all() and (layout[13:] == 0).all()
    assert (grab[1:13] == 1).all() and grab[32] == 4 and grab[30] == 16
    assert all(grab[n] == 8 for n in range(13, 32) if n != 30)
    assert np.all(grab >= 1)
    pad = api.lds_pad_get()
    assert pad.tolist() == [1 if n in (4, 8, 12, 16, 20, 24, 32) else 0 for n in range(33)]


def test_apply_table_with_padding(tmp_path):
    from matrixinversion_amd import api, autotune

    layout, grab = api.tuning_get()
    pad = api.lds_pad_get()
    try:
        t = {"layout": layout.tolist(), "grab": grab.tolist(), "pad": [0] * 33}
        t["pad"][20] = 1
        autotune.apply_table(t)
        assert api.lds_pad_get()[20] == 1 and api.lds_pad_get()[32] == 0
    finally:
        api.lds_pad_set(pad)
        api.tuning_set(layout, grab)
