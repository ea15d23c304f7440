"""The Gram tile autotuner's host logic: keys, choice, cache file (no GPU: the launches
and the timer are stand-ins)."""
import json

from vtem_amd.autotune import TILES, GramTuner, default_tile, tune_key


def _tuner(path, times, enabled=True):
    def timer(run, tile, reps):
        run(tile)
        return times[tile]

    return GramTuner(path=str(path), enabled=enabled, reps=1, timer=timer), None


def test_keys_bucket_workgroups_and_neurons():
    # Ds 75 -> 5 blocks of 16 -> 15 workgroups per problem
    assert tune_key("bf16x3", 17, 75, 1, 150) == "bf16x3/dt17/wg4/n8"
    assert tune_key("bf16x3", 17, 75, 2, 150) == "bf16x3/dt17/wg5/n8"
    assert tune_key("fp16", 9, 80, 1, 256) == "fp16/dt9/wg4/n8"
    assert tune_key("fp16", 9, 80, 1, 257) == "fp16/dt9/wg4/n9"
    assert default_tile("bf16x3") == 2 and default_tile("fp16") == 4


def test_tuner_picks_fastest_and_caches(tmp_path):
    path = tmp_path / "tiles.json"
    t, _ = _tuner(path, {2: 1.0, 4: 0.7, 8: 0.9})
    ran = []
    assert t.launch(ran.append, "bf16x3", 17, 600, 4, 2000) == 4
    assert ran == list(TILES)                     # each tiling timed once on the launch
    key = tune_key("bf16x3", 17, 600, 4, 2000)
    assert t.timings[key] == {2: 1.0, 4: 0.7, 8: 0.9}
    assert json.loads(path.read_text())["choices"][key] == 4
    # a second launch with the key runs the cached tiling once, untimed
    ran.clear()
    assert t.launch(ran.append, "bf16x3", 17, 600, 4, 2000) == 4
    assert ran == [4]
    # a new process reads the file
    def no_timer(run, tile, reps):
        raise AssertionError("cached keys are not timed")
    t2 = GramTuner(path=str(path), enabled=True, timer=no_timer)
    ran.clear()
    assert t2.launch(ran.append, "bf16x3", 17, 600, 4, 2000) == 4 and ran == [4]


def test_tuner_ties_prefer_default_and_disabled_uses_default(tmp_path):
    t, _ = _tuner(tmp_path / "a.json", {2: 0.5, 4: 0.5, 8: 0.5})
    assert t.launch(lambda tile: None, "fp16", 13, 300, 1, 500) == 4
    t, _ = _tuner(tmp_path / "b.json", {2: 0.1, 4: 0.5, 8: 0.5}, enabled=False)
    ran = []
    assert t.launch(ran.append, "fp16", 13, 300, 1, 500) == 4 and ran == [4]
    assert not (tmp_path / "b.json").exists()


def test_tuner_ignores_bad_cache_files(tmp_path):
    path = tmp_path / "tiles.json"
    path.write_text("{not json")
    t, _ = _tuner(path, {2: 0.3, 4: 0.2, 8: 0.1})
    assert t.choices == {}
    assert t.launch(lambda tile: None, "bf16x3", 5, 100, 1, 100) == 8
    path.write_text(json.dumps({"choices": {"bf16x3/dt5/wg5/n7": 3}}))   # not a tiling
    assert GramTuner(path=str(path)).choices == {}
    # an unwritable cache path keeps the choice in memory
    (tmp_path / "plain_file").write_text("")
    t, _ = _tuner(tmp_path / "plain_file" / "tiles.json", {2: 0.3, 4: 0.2, 8: 0.1})
    assert t.launch(lambda tile: None, "bf16x3", 5, 100, 1, 100) == 8
