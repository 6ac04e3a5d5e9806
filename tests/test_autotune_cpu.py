"""CPU checks of the launch-shape tuning: the library keeps one shape per
host thread and validates it, and the autotuner's persisted cache (shape
normalisation, keys, file round trip, the environment switch) works without
a GPU."""
import ctypes
import json
import threading

import pytest

from torchcompactradius_amd import _lib, autotune


def _get():
    t = _lib.TcrTuning()
    _lib.check(_lib.load().tcr_get_tuning(ctypes.byref(t)))
    return (t.cell_workgroups, t.stencil_workgroups, t.cell_threads, t.cells_per_wave)


def test_tuning_is_per_thread():
    L = _lib.load()
    _lib.check(L.tcr_set_tuning(ctypes.byref(_lib.TcrTuning(1024, 512, 128, 4))))
    seen = {}

    def other():
        seen["start"] = _get()
        _lib.check(L.tcr_set_tuning(ctypes.byref(_lib.TcrTuning(8, 0, 64, 2))))
        seen["set"] = _get()

    th = threading.Thread(target=other)
    th.start()
    th.join()
    assert seen["start"] == (0, 0, 0, 0)          # a new thread starts from the defaults
    assert seen["set"] == (8, 0, 64, 2)
    assert _get() == (1024, 512, 128, 4)          # untouched by the other thread
    _lib.check(L.tcr_set_tuning(None))
    assert _get() == (0, 0, 0, 0)


@pytest.mark.parametrize("bad", [(-1, 0, 0, 0), (0, 1 << 21, 0, 0), (0, 0, 96, 0), (0, 0, 512, 0), (0, 0, 0, 65),
                                 (0, 0, 0, -1)])
def test_invalid_tuning_rejected(bad):
    L = _lib.load()
    assert L.tcr_set_tuning(ctypes.byref(_lib.TcrTuning(*bad))) == _lib.TCR_ERR_ARG
    assert _get() == (0, 0, 0, 0)


def test_install_tracks_the_calling_thread():
    autotune.install((256, 0))
    assert _get() == (256, 0, 0, 0)
    out = {}

    def other():
        autotune.install((256, 0))                # not a no-op here: this thread never set it
        out["v"] = _get()
        autotune.install(None)

    th = threading.Thread(target=other)
    th.start()
    th.join()
    assert out["v"] == (256, 0, 0, 0)
    autotune.install(None)
    assert _get() == (0, 0, 0, 0)


def test_normalize_and_key():
    assert autotune.normalize(None) is None
    assert autotune.normalize((8, 64)) == (8, 64, 0, 0)
    assert autotune.normalize([1, 2, 128]) == (1, 2, 128, 0)
    with pytest.raises(ValueError):
        autotune.normalize((1,))

    class P:
        key = (3, 1000, 1000, ((3, 1), (2, 0), (3, 1)))
    k = autotune.shape_key(P(), "AMD Instinct MI355X")
    assert json.loads(k) == ["AMD Instinct MI355X", [3, 1000, 1000, [[3, 1], [2, 0], [3, 1]]]]


def test_cache_file_round_trip(tmp_path, monkeypatch):
    path = str(tmp_path / "sub" / "autotune.json")
    monkeypatch.setenv("TCR_AUTOTUNE_CACHE", path)
    assert autotune.cache_path() == path
    assert autotune.load_file() == {}
    autotune.save_file({"a": (1024, 0, 128, 4)})
    autotune.save_file({"b": (0, 512, 0, 0)})       # merged with what is there
    assert autotune.load_file() == {"a": (1024, 0, 128, 4), "b": (0, 512, 0, 0)}
    with open(path, "w") as f:
        f.write("{not json")
    assert autotune.load_file() == {}               # unreadable: ignored
    monkeypatch.setenv("TCR_AUTOTUNE_CACHE", "")
    assert autotune.cache_path() is None
    autotune.save_file({"c": (1, 1, 0, 0)})         # disabled: nothing written
    assert autotune.load_file() == {}
