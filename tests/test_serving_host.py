"""Host side of the serving loop and of admission by footprint (no GPU)."""
import numpy as np


def test_serving_mix_is_seeded_and_supported(oaxis_mod):
    oaxis = oaxis_mod
    from oaxis.serve import serving_mix
    a = serving_mix(5, 40, rate=20.0)
    b = serving_mix(5, 40, rate=20.0)
    assert [(r.kind, r.arrival, r.steps, r.text) for r in a] == [(r.kind, r.arrival, r.steps, r.text) for r in b]
    arr = [r.arrival for r in a]
    assert arr == sorted(arr) and arr[0] > 0.0
    for r in a:
        p = oaxis.params_from_text(r.text)
        assert oaxis.lib().oa_params_check_supported(p, None, 0) == 0, r.text
        assert 200 <= r.steps <= 2000
    kinds = [r.kind for r in serving_mix(1, 2000, rate=1.0)]
    frac = {k: kinds.count(k) / len(kinds) for k in set(kinds)}
    assert abs(frac["schwarzschild"] - 0.4) < 0.05
    assert abs(frac["brill"] - 0.3) < 0.05
    assert abs(frac["scalar_pulse"] - 0.3) < 0.05


def test_footprint_counts_every_allocation(oaxis_mod):
    oaxis = oaxis_mod
    from oaxis import workloads
    p1 = oaxis.params_from_text(workloads.schwarzschild(1.0, 64, 128, 0))
    p2 = oaxis.params_from_text(workloads.schwarzschild(1.0, 64, 128, 1))
    p3 = oaxis.params_from_text(workloads.scalar_pulse(0, 0))
    f1, f2, f3 = oaxis.footprint(p1), oaxis.footprint(p2), oaxis.footprint(p3)
    # one level of 64 x 128, 12 fields: 4 state copies of (64+3) x ld planes, ld = 144
    plane = 67 * 144 * 8
    assert f1 > 4 * 12 * plane and f1 < 6 * 12 * plane
    assert f2 > 2 * f1 * 0.95
    assert f3 > 0 and np.isfinite(f3)
