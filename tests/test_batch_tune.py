"""The batched pipeline's tuning table persists across processes (CPU)."""

import json
import os
import subprocess
import sys
from pathlib import Path

from cucim_amd import batch


def _fresh(monkeypatch, path):
    monkeypatch.setenv("CUCIM_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(batch, "_TUNED", {})
    monkeypatch.setattr(batch, "_TUNED_LOADED", False)


def test_save_then_load_in_another_process(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    _fresh(monkeypatch, path)
    batch._TUNED["gray_gauss_sobel|K=13|x"] = 32
    batch.tune_save()
    data = json.loads(path.read_text())
    assert data["entries"] == {"gray_gauss_sobel|K=13|x": 32}
    code = ("from cucim_amd import batch; "
            "print(batch._tuned('gray_gauss_sobel|K=13|x', 16), batch._tuned('other', 16))")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       env={**os.environ, "CUCIM_AMD_TUNE_CACHE": str(path)},
                       cwd=str(Path(__file__).resolve().parents[1]))
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["32", "16"]


def test_merge_keeps_other_processes_entries(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"version": 1, "entries": {"a": 16, "b": 32}}))
    _fresh(monkeypatch, path)
    assert batch._tuned("a", 0) == 16
    batch._TUNED["b"] = 16  # chosen here: wins over the file
    batch._TUNED["c"] = 32
    batch.tune_save()
    assert json.loads(path.read_text())["entries"] == {"a": 16, "b": 16, "c": 32}


def test_bad_or_missing_file_is_ignored(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    _fresh(monkeypatch, path)
    assert batch._tuned("a", 16) == 16
    path.write_text(json.dumps({"entries": {"a": "32", "b": 32}}))
    _fresh(monkeypatch, path)
    assert batch._tuned("a", 16) == 16 and batch._tuned("b", 16) == 32
    _fresh(monkeypatch, tmp_path / "missing" / "tune.json")
    assert batch._tuned("a", 16) == 16


def test_disabled(tmp_path, monkeypatch):
    _fresh(monkeypatch, "off")
    batch._TUNED["a"] = 32
    batch.tune_save()  # writes nothing
    assert batch._tuned("a", 16) == 32
    assert not list(tmp_path.iterdir())
