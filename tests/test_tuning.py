"""Pass settings per size bucket (gls_ctx_get_tuning / gls_ctx_set_tuning),
on a host-only context: defaults, buckets by floor(log2(objects)), every
bucket at once, and refusal of values out of range."""
import pytest

import culs_amd

DEFAULTS = {"cut_wide": 2048, "rw_sweep_width": 1024, "strash_table_factor": 4, "strash_wave_batch": 4,
            "balance_narrow_max": 4096, "cut_levels": 3}


def test_defaults_and_buckets():
    c = culs_amd.Context(-1)
    for n in (1, 1000, 1 << 20, (1 << 31) - 1):
        assert c.get_tuning(n) == DEFAULTS
    c.set_tuning(1 << 20, cut_wide=512, strash_wave_batch=16)
    got = c.get_tuning(1 << 20)
    assert got["cut_wide"] == 512 and got["strash_wave_batch"] == 16
    # the same bucket: [2^20, 2^21)
    assert c.get_tuning((1 << 21) - 1) == got
    # neighbouring buckets keep the defaults
    assert c.get_tuning((1 << 20) - 1) == DEFAULTS
    assert c.get_tuning(1 << 21) == DEFAULTS


def test_every_bucket():
    c = culs_amd.Context(-1)
    c.set_tuning(-1, rw_sweep_width=4096, strash_table_factor=8)
    for n in (1, 77, 1 << 30):
        t = c.get_tuning(n)
        assert t["rw_sweep_width"] == 4096 and t["strash_table_factor"] == 8


@pytest.mark.parametrize("bad", [{"strash_table_factor": 3}, {"strash_table_factor": 16}, {"cut_wide": 0},
                                 {"strash_wave_batch": 0}, {"strash_wave_batch": 65}, {"balance_narrow_max": 0},
                                 {"rw_sweep_width": -5}, {"cut_levels": 0}, {"cut_levels": 7}])
def test_out_of_range_refused(bad):
    c = culs_amd.Context(-1)
    with pytest.raises(culs_amd.GlsError):
        c.set_tuning(1000, **bad)
    assert c.get_tuning(1000) == DEFAULTS


def test_autotune_needs_a_device():
    c = culs_amd.Context(-1)
    a = c.gen_random(16, 200, 8, 50, 1)
    with pytest.raises(culs_amd.GlsError):
        c.autotune([a])
