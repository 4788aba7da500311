"""On-disk snapshot format, checked on the CPU against the text the
reference's writer emits (python/csrc/native/src/algorithm/data_recorder.cu:
157-205 metadata.json, 260-273 snapshot.json, 68-76 directory names,
129-136 save schedule) with synthetic arrays."""
import json

import numpy as np
import pytest

from cuda_moea.record import SnapshotRecorder, generation_directory, load_snapshot

# The reference writer's output for these values, line for line.
METADATA = """{
  "format_version": 1,
  "scalar_type": "float32",
  "index_type": "int32",
  "algorithm_name": "NSGA-III",
  "problem_name": "DTLZ1",
  "population_size": 6,
  "dimension": 4,
  "objective_count": 3,
  "constraint_count": 0,
  "max_generations": 10,
  "random_seed": 2887,
  "generation_interval": 4,
  "reference_direction_count": 5,
  "layouts": {
    "variables": "row-major (N,D)",
    "objectives": "objective-major (M,N)",
    "constraints": "(N)",
    "reference_directions": "objective-major (M,K)",
    "bounds": "separate lower/upper arrays (D)"
  }
}
"""

SNAPSHOT = """{
  "generation": 8,
  "population_size": 6,
  "dimension": 4,
  "objective_count": 3,
  "active_count": 6,
  "reference_direction_count": 5,
  "auxiliary_index_count": 6
}
"""


def _recorder(root, name="DTLZ1", interval=4, directions=True):
    rec = SnapshotRecorder(root, interval)
    w = np.arange(15, dtype=np.float32).reshape(3, 5) / 7 if directions else None
    rec.initialize(algorithm_name="NSGA-III", problem_name=name, population_size=6, dimension=4,
                   objective_count=3, constraint_count=0, max_generations=10, random_seed=2887,
                   lower_bounds=[0.0, -1.0, -2.0, -3.0], upper_bounds=[1.0, 2.0, 3.0, 4.0],
                   initial_directions=w)
    return rec, w


def test_metadata_and_snapshot_text_match_the_reference_writer(tmp_path):
    rec, w = _recorder(tmp_path)
    assert (tmp_path / "metadata.json").read_bytes() == METADATA.encode()
    assert np.fromfile(tmp_path / "lower_bounds.bin", "<f4").tolist() == [0.0, -1.0, -2.0, -3.0]
    assert (tmp_path / "upper_bounds.bin").stat().st_size == 16
    assert (tmp_path / "initial_reference_directions.bin").read_bytes() == w.astype("<f4").tobytes()
    rng = np.random.default_rng(0)
    x = rng.random((6, 4), dtype=np.float32)
    f = rng.random((3, 6), dtype=np.float32)
    cv = np.zeros(6, np.float32)
    fr = np.array([0, 1, 0, 2, 1, 0], np.int32)
    out = rec.save(8, x, f, cv, w, fr, 6)
    assert out == tmp_path / "generation_000008"
    assert (out / "snapshot.json").read_bytes() == SNAPSHOT.encode()
    assert (out / "variables.bin").read_bytes() == x.tobytes()          # row-major (N, D)
    assert (out / "objectives.bin").read_bytes() == f.tobytes()         # objective-major (M, N)
    assert (out / "selector_auxiliary_indices.bin").read_bytes() == fr.astype("<i4").tobytes()
    back = load_snapshot(tmp_path, 8)
    np.testing.assert_array_equal(back["variables"], x)
    np.testing.assert_array_equal(back["objectives"], f)
    np.testing.assert_array_equal(back["auxiliary_indices"], fr)
    np.testing.assert_array_equal(back["reference_directions"], w)
    assert back["metadata"]["layouts"]["objectives"] == "objective-major (M,N)"


def test_optional_files_and_counts(tmp_path):
    rec, _ = _recorder(tmp_path, directions=False)
    assert not (tmp_path / "initial_reference_directions.bin").exists()
    assert json.loads((tmp_path / "metadata.json").read_text())["reference_direction_count"] == 0
    out = rec.save(3, np.ones((6, 4), np.float32), np.ones((3, 6), np.float32), np.ones(6, np.float32),
                   None, np.zeros(0, np.int32), 4)
    snap = json.loads((out / "snapshot.json").read_text())
    assert snap["auxiliary_index_count"] == 0 and snap["reference_direction_count"] == 0
    assert snap["active_count"] == 4
    assert not (out / "selector_auxiliary_indices.bin").exists()
    assert not (out / "reference_directions.bin").exists()


@pytest.mark.parametrize("name,text", [
    ('a"b\\c', '"a\\"b\\\\c"'),
    ("tab\tnl\nret\rbs\bff\f", '"tab\\tnl\\nret\\rbs\\bff\\f"'),
    ("ctl\x01\x1f", '"ctl\\u0001\\u001f"'),
    ("Zürich/π", '"Zürich/π"'),                 # bytes >= 0x80 and '/' pass through
    ("", "null"),                               # an empty name is written as null
])
def test_problem_name_escaping(tmp_path, name, text):
    _recorder(tmp_path, name=name)
    raw = (tmp_path / "metadata.json").read_text(encoding="utf-8")
    assert f'  "problem_name": {text},\n' in raw


def test_save_schedule_and_directory_names(tmp_path):
    rec = SnapshotRecorder(tmp_path, 4)
    assert [g for g in range(11) if rec.should_save(g, 10)] == [0, 4, 8, 10]
    assert [rec.next_save(g, 10) for g in (0, 3, 4, 8, 9)] == [4, 4, 8, 10, 10]
    assert SnapshotRecorder(tmp_path, 1).next_save(5, 10) == 6
    assert generation_directory(tmp_path, 7).name == "generation_000007"
    assert generation_directory(tmp_path, 1234567).name == "generation_1234567"
    with pytest.raises(ValueError, match="interval must be positive"):
        SnapshotRecorder(tmp_path, 0)
    with pytest.raises(RuntimeError, match="initialized"):
        rec.save(0, np.ones((2, 2), np.float32), np.ones((1, 2), np.float32), np.ones(2, np.float32),
                 None, None, 2)


def test_algorithm_checks_the_interval_without_a_gpu(tmp_path):
    import cuda_moea as cm
    with pytest.raises(ValueError, match="interval must be positive"):
        cm.NSGA3(population_size=92, max_generations=3, save_interval=0)
    alg = cm.NSGA3(population_size=92, max_generations=3, save_data="", save_interval=2)
    assert str(alg._recorder.root) == "cuda_moea_data" and alg._recorder.interval == 2
    assert cm.NSGA3(population_size=92, max_generations=3)._recorder is None
