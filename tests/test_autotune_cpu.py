"""Launch-tuner candidate sets and flag encoding (CPU; no launches)."""

from tests.conftest import load_package


def test_candidates_and_flags():
    at = load_package("autotune")
    lib = load_package("_lib")
    aligned = at.candidates(128)
    assert aligned and not any(c[4] for c in aligned)  # whole-wavefront rows: (z, y) tiles only
    assert all(bz * by <= 256 and (bz * by) % 64 == 0 for bz, by, *_ in aligned)
    assert not any(bz > 64 and 128 < bz // 2 for bz, by, *_ in aligned)  # no mostly idle row tiles
    assert any(bz == 256 for bz, *_ in aligned)  # half-idle 256-wide tiles are timed (fastest at 128^3)
    assert not any(c[0] == 256 and not c[4] for c in at.candidates(96))  # (z, y) tiles; flat ones are dense
    odd = at.candidates(136)
    flat = [c for c in odd if c[4]]
    assert {c[0] for c in flat} == set(at.FLAT) and all(c[1] == 1 for c in flat)
    assert len(flat) == len(at.FLAT) * len(at.NT) * len(at.BPC)
    # moving spheres: the flat mapping is searched at aligned rows too
    moving = at.candidates(512, moving=True)
    assert {c for c in moving if not c[4]} == set(at.candidates(512))
    assert len([c for c in moving if c[4]]) == len(at.FLAT) * len(at.NT) * len(at.BPC)
    # flag layout of psc_step: bit 0 non-temporal, bits 8-15 cap, 16-17 cells per thread, bit 20 flat
    assert at.flags_of(1, 6) == 0x10601
    assert at.flags_of(0, 3, 2) == 0x20300
    assert at.flags_of(1, 4, 1, True) == lib.PSC_STEP_NONTEMPORAL | 0x400 | 0x10000 | lib.PSC_STEP_LINEAR
    # the in-place bit belongs to psc_run_steps plans, never to a psc_step flag word
    for c in odd:
        assert at.flags_of(c[2], c[3], 1, c[4]) & lib.PSC_RUN_INPLACE == 0
    assert lib.PSC_RUN_INPLACE & (lib.PSC_STEP_LINEAR | 0x3ffff) == 0
