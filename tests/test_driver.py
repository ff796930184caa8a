"""Host logic of the monitored multi-step driver: stop schedule, report arithmetic and the
checkpoint writer (no GPU)."""
import numpy as np
import pytest

from bia_3d_gpu_amd import driver, io


def test_schedule_merges_monitor_and_checkpoint_steps():
    assert driver.schedule(7, 2, 3) == [(2, True, False), (3, False, True), (4, True, False), (6, True, True),
                                        (7, True, False)]
    assert driver.schedule(4, 0, 0) == [(4, True, False)]
    assert driver.schedule(4, 4, 2) == [(2, False, True), (4, True, True)]
    assert driver.schedule(0, 3, 3) == []
    with pytest.raises(ValueError):
        driver.schedule(-1)


def test_report_arithmetic():
    inv = np.zeros((3, 2, 11))
    inv[:, :, 0] = [[1.0, 2.0], [1.1, 2.0], [1.2, 2.5]]    # KE
    inv[:, :, 1] = [[-2.0, -4.0], [-2.1, -4.0], [-2.2, -4.0]]  # PE
    inv[:, :, 5] = [[0.0, 1.0], [1e-15, 1.0], [2e-15, 1.0]]  # Ua
    stats = np.zeros((3, 16), dtype=np.uint64)
    stats[1, 0], stats[1, 2], stats[1, 4], stats[1, 6] = 10, 5, 20, 10
    stats[2, 0], stats[2, 2], stats[2, 4], stats[2, 6], stats[2, 8 + 2] = 20, 10, 45, 20, 1
    r = driver.Report(sizes=[4, 5], steps=[0, 2, 4], invariants=inv, stats=stats)
    np.testing.assert_allclose(r.energy_error(), [[0, 0], [0, 0], [0, 0.25]], atol=1e-15)
    assert r.momentum_error()[2, 0, 0] == 2e-15
    np.testing.assert_array_equal(r.interval_stats()[:, [0, 2, 4, 6, 10]], [[10, 5, 20, 10, 0], [10, 5, 25, 10, 1]])
    h = r.histogram()
    assert h["pairs"] == 30 and h["hyperbola"] == 20 and h["ellipse"] == 10 and h["iteration cap"] == 1
    assert h["evaluations_per_pair"] == pytest.approx(65 / 30)
    assert r.histogram(r.interval_stats()[0])["evaluations_per_pair"] == pytest.approx(2.0)


def test_checkpoint_writer_writes_one_record_per_system(tmp_path):
    import torch

    class Done:
        def synchronize(self):
            pass

    rng = np.random.default_rng(3)
    sizes = [3, 5, 1]
    st = rng.standard_normal((6, 10))  # one spare column beyond the systems
    w = driver._CheckpointWriter()
    path = str(tmp_path / "ck.bbb")
    w.submit(path, torch.from_numpy(st), sizes, Done())
    w.join()
    recs = io.read_bbb(path)
    assert [r.shape for r in recs] == [(6, 3), (6, 5), (6, 1)]
    assert np.array_equal(np.concatenate(recs, axis=1), st[:, :9])
    w.submit(str(tmp_path / "missing_dir" / "ck.bbb"), torch.from_numpy(st), sizes, Done())
    with pytest.raises(OSError):
        w.join()


def test_report_without_statistics():
    inv = np.zeros((2, 1, 11))
    inv[:, 0, 0] = [1.0, 1.0]
    inv[:, 0, 1] = [-2.0, -2.0]
    r = driver.Report(sizes=[3], steps=[0, 5], invariants=inv, stats=None)
    assert np.all(r.energy_error() == 0)
    with pytest.raises(ValueError):
        r.histogram()
    with pytest.raises(ValueError):
        r.interval_stats()
