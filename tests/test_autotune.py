"""Host side of the launch-shape autotuner (no GPU needed)."""

import json

from fresnel_amd import autotune


class _FakeDev:
    def __init__(self):
        self.knobs = {'stack_lds': 8, 'batch_paths': 1 << 25, 'tail_rays': 32768, 'trav_chunk': 256, 'trav_refill': 65, 'trav_packet': 2, 'primary_tiles': 2,
                      'tile_group': 4, 'primary_chunk': 1024, 'fuse_primary': 2, 'fuse_bounce': 1, 'inflight': 2, 'xcd_bands': 2}

    def describe(self):
        return "Enabled HIP devices:\n[0]: AMD Instinct MI355X, 256 CUs\n"

    def tune(self, key, value):
        prev = self.knobs[key]
        if value > 0:
            self.knobs[key] = value
        return prev


class _Device:
    def __init__(self):
        self._device = _FakeDev()


def test_arch_and_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv('FRESNEL_AMD_TUNING', str(tmp_path / 'tuning.json'))
    dev = _Device()
    assert autotune.arch(dev) == 'AMD Instinct MI355X, 256 CUs'
    assert autotune.load(dev, 'k') is None
    autotune.store(dev, 'k', {'stack_lds': 16})
    assert autotune.load(dev, 'k') == {'stack_lds': 16}
    data = json.loads((tmp_path / 'tuning.json').read_text())
    assert data == {'AMD Instinct MI355X, 256 CUs': {'k': {'stack_lds': 16}}}


def test_apply_and_current():
    dev = _Device()
    prev = autotune.apply(dev, {'stack_lds': 16, 'tail_rays': 100})
    assert prev == {'stack_lds': 8, 'tail_rays': 32768}
    assert autotune.current(dev) == {'stack_lds': 16, 'batch_paths': 1 << 25, 'tail_rays': 100, 'trav_chunk': 256, 'trav_refill': 65, 'trav_packet': 2, 'primary_tiles': 2,
                                      'tile_group': 4, 'primary_chunk': 1024, 'fuse_primary': 2, 'fuse_bounce': 1, 'inflight': 2, 'xcd_bands': 2}
