"""Target-SRO layout, counts and synthetic targets (pyhea/io/input.py)."""
import os

import numpy as np
import pytest

from pyhea_amd.config import input_config, read_target_sro
from pyhea_amd.problem import largest_remainder, random_target, target_from_lower_triangle


def test_k2_expansion_reference_fixture():
    # pyhea tests/io/test_config.py:106-111 and tests/io/test_input.py:134-139
    assert np.array_equal(target_from_lower_triangle([[0.1, 0.2, 0.3]], 2),
                          [[0.1, 0.2, 0.2, 0.3], [0] * 4, [0] * 4])
    assert np.array_equal(
        target_from_lower_triangle([[0.0, 0.1, 0.2], [0.3, 0.4, 0.5], [0.6, 0.7, 0.8]], 2),
        [[0.0, 0.1, 0.1, 0.2], [0.3, 0.4, 0.4, 0.5], [0.6, 0.7, 0.7, 0.8]])


def test_k3_k4_mapping_table():
    # the explicit index table of pyhea/io/input.py:430-455
    tri3 = np.arange(6.0)
    assert target_from_lower_triangle([tri3], 3)[0].tolist() == [0, 1, 3, 1, 2, 4, 3, 4, 5]
    tri4 = np.arange(10.0)
    assert target_from_lower_triangle([tri4], 4)[0].tolist() == \
        [0, 1, 3, 6, 1, 2, 4, 7, 3, 4, 5, 8, 6, 7, 8, 9]


def test_k6_supported():
    full = target_from_lower_triangle([np.arange(21.0)], 6)[0].reshape(6, 6)
    assert np.array_equal(full, full.T) and full[5, 5] == 20


def test_wrong_length_raises():
    with pytest.raises(ValueError):
        target_from_lower_triangle([[0.1, 0.2]], 2)


def test_example_target_file(tmp_path):
    p = tmp_path / "TARGET_SRO"
    p.write_text("# comment\nFirst Shell\n-0.038\n-0.008  0.067\n 0.351 -0.268 -0.717\n"
                 "-0.167  0.034  0.388  0.146\n")
    full = read_target_sro(str(p), 4)
    assert full.shape == (3, 16)
    assert full[0, 3] == -0.167 and full[0, 12] == -0.167 and full[0, 15] == 0.146
    assert not full[1:].any()


def test_largest_remainder():
    assert largest_remainder([0.30, 0.25, 0.20, 0.15, 0.10], 8192) == [2458, 2048, 1638, 1229, 819]
    assert largest_remainder([0.25] * 4, 4000) == [1000] * 4
    assert sum(largest_remainder([0.999, 0.0005, 0.0005], 10)) == 10
    assert min(largest_remainder([0.999, 0.0005, 0.0005], 10)) >= 1


def test_random_target_sum_rule():
    counts = [2458, 2048, 1638, 1229, 819]
    t = random_target(counts, 3, 7)
    c = np.asarray(counts) / sum(counts)
    for s in range(3):
        a = t[s].reshape(5, 5)
        assert np.allclose(a, a.T) and np.allclose(a @ c, 0.0, atol=1e-14)
        off = a[~np.eye(5, dtype=bool)]
        assert off.min() >= -0.15 and off.max() <= 0.05


def test_input_config_yaml(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    (tmp_path / "TARGET_SRO").write_text("Shell\n0.1\n0.2 0.3\n")
    cfg = tmp_path / "config.yaml"
    cfg.write_text("type: 2\nelement: [16, 16]\ncell_dim: [2, 2, 2]\ndevice: gpu\n"
                   "solutions: 4\ntotal_iter: 2\nparallel_task: 8\nconverge_depth: 2\n"
                   "weight: [4.0, 1.0]\nstructure: FCC\ntarget_sro: TARGET_SRO\n")
    c = input_config(str(cfg))
    assert c.type == 2 and c.element == [16, 16] and c.device == "gpu"
    assert c.weight == [4.0, 1.0] and c.parallel_task == 8
    assert np.array_equal(c.target_sro[0], [0.1, 0.2, 0.2, 0.3])


def test_target_sro_path_is_relative_to_the_working_directory(tmp_path, monkeypatch):
    """A relative target_sro names a file in the working directory, not beside
    the YAML file (pyhea/io/input.py:373-374)."""
    ydir, wdir = tmp_path / "conf", tmp_path / "work"
    ydir.mkdir()
    wdir.mkdir()
    (ydir / "TARGET_SRO").write_text("Shell\n0.1\n0.2 0.3\n")
    (wdir / "TARGET_SRO").write_text("Shell\n0.4\n0.5 0.6\n")
    cfg = ydir / "config.yaml"
    cfg.write_text("type: 2\nelement: [16, 16]\ntarget_sro: TARGET_SRO\n")
    monkeypatch.chdir(wdir)
    assert np.array_equal(input_config(str(cfg)).target_sro[0], [0.4, 0.5, 0.5, 0.6])
    (wdir / "TARGET_SRO").unlink()
    with pytest.raises(FileNotFoundError):
        input_config(str(cfg)).target_sro


def test_input_config_rejects_unknown_key(tmp_path):
    cfg = tmp_path / "c.yaml"
    cfg.write_text("type: 2\nelement: [16, 16]\nbogus: 1\n")
    with pytest.raises(ValueError):
        input_config(str(cfg))
