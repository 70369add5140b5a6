"""The autotuner's host logic: candidate launch shapes from the occupancy
model, choice of the fastest, and the cache (timing replaced by a model)."""
import types

from fvm2d_amd import autotune
from fvm2d_amd.solver import Inputs


def test_candidates_fill_whole_rounds():
    for np_ in (1_000, 4_224, 66_048, 263_168, 1_050_624):
        c = autotune.candidates(np_)
        assert c == sorted(set(c)) and all(24 <= ps <= 128 for ps in c)
        if np_ >= 64 * 1024:
            # some candidate fills its last round of 1024 resident workgroups
            def fill(ps):
                n = -(-np_ // ps)
                return n / (1024 * -(-n // 1024))
            assert max(fill(ps) for ps in c) > 0.95
    assert 65 in autotune.candidates(66_048)           # 1017 patches: one round of 1024


def test_tune_picks_fastest_and_caches(tmp_path):
    geo = types.SimpleNamespace(dims=dict(np=66_048, nt=131_072))
    inp = Inputs(iflow=1, precision=0)
    calls = []

    def timer(threads, ps, xcd):
        calls.append((threads, ps, xcd))
        return 100.0 + abs(ps - 65) + (0.5 if xcd == 0 else 0.0) + (3.0 if threads == 512 else 0.0)

    path = str(tmp_path / "tune.json")
    c = autotune.tune(None, geo, inp, cache_path=path, device="gfx950", timer=timer)
    assert (c.threads, c.patch_size, c.xcd) == (256, 65, 1) and c.us_per_iter == 100.0
    assert len(calls) == 2 * len(autotune.shapes(66_048, inp))
    assert {t for t, _, _ in calls} == {256, 512}          # Euler, limited: both sizes
    n = len(calls)
    again = autotune.tune(None, geo, inp, cache_path=path, device="gfx950", timer=timer)
    assert len(calls) == n and again.patch_size == 65          # from the cache
    other = autotune.tune(None, geo, Inputs(iflow=1, precision=1), cache_path=path,
                          device="gfx950", timer=timer)
    assert len(calls) == 2 * n and other.patch_size == 65      # another key: timed again


def test_block_sizes_per_kind():
    """512-thread workgroups exist for limited Euler solves (every flux)
    only; their patch candidates are about twice as large and fill whole
    rounds of two workgroups a CU."""
    for iflux in (1, 2, 3):
        assert autotune.block_sizes(Inputs(iflow=1, iflux=iflux, ilimit=1)) == [256, 512]
    for kw in (dict(iflow=2), dict(iflow=3), dict(iflow=1, ilimit=0), dict(iflow=1, iflux=2, ilimit=0)):
        assert autotune.block_sizes(Inputs(**kw)) == [256]
    sh = autotune.shapes(66_048, Inputs(iflow=1))
    big = [ps for t, ps in sh if t == 512]
    assert big and all(48 <= ps <= 256 for ps in big) and 129 in big      # 512 patches: one round
    small = [ps for t, ps in sh if t == 256]
    assert small == autotune.candidates(66_048)
