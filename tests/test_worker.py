"""The batched GPU feature worker (features.py:270-381 protocol): tasks in, ("gpu" | "gpu_off",
wid, msg), ("ok", (idx, n, err, timings)) per photo and ("exit", "gpuN") out."""
import queue

import numpy as np
import pytest

from test_gpu_features import per_image_path


def _tasks(tmp_path, shapes, nfeat=1500):
    from fieldstitch_amd import synth
    rng = np.random.default_rng(8)
    tasks = []
    for k, (h, w) in enumerate(shapes):
        mos = synth.Mosaic(seed=100 + k)
        mos.polys = synth.random_polygons(rng, 10, 0, 0, w, h)
        g = mos.render(np.eye(3)[:2], h, w, rng)
        p = tmp_path / ("img%d.npy" % k)
        np.save(p, g)
        tasks.append((k, str(p), str(tmp_path / ("%d.npz" % k)), None, 1, nfeat))
    return tasks


def _run(tasks, n_threads=2, batch=4, **kw):
    from fieldstitch_amd import features
    tq, rq = queue.Queue(), queue.Queue()
    for t in tasks:
        tq.put(t)
    for _ in range(n_threads):
        tq.put(None)
    features.gpu_worker_main(tq, rq, n_threads, 0, batch=batch, **kw)
    out = []
    while not rq.empty():
        out.append(rq.get())
    return out


def test_worker_without_gpu_reports_every_photo(tmp_path, monkeypatch):
    from fieldstitch_amd import features

    def no_gpu():
        raise RuntimeError("no ROCm GPU visible")
    monkeypatch.setattr(features, "_gpu_ready", no_gpu)
    tasks = _tasks(tmp_path, [(60, 80), (64, 96)])
    (tmp_path / "notaphoto.jpg").write_bytes(b"not an image")
    tasks.append((7, str(tmp_path / "notaphoto.jpg"), str(tmp_path / "7.npz"), None, 1, 1500))
    msgs = _run(tasks)
    assert msgs[0][0] == "gpu_off" and "no ROCm GPU" in msgs[0][2]
    assert msgs[-1] == ("exit", "gpu0")
    ok = {m[1][0]: m[1] for m in msgs if m[0] == "ok"}
    assert sorted(ok) == [0, 1, 7]
    assert all(r[1] == 0 and r[2] for r in ok.values())
    assert "cannot identify" in ok[7][2] and "GPU unavailable" in ok[0][2]


@pytest.mark.gpu
def test_worker_matches_sift_gpu(tmp_path, gpu):
    from fieldstitch_amd import features
    shapes = [(240, 320), (240, 320), (300, 410), (240, 320), (180, 200)]
    tasks = _tasks(tmp_path, shapes)
    tasks[3] = tasks[3][:5] + (800,)                        # a second nfeat group
    tasks.append((9, str(tmp_path / "missing.npy"), str(tmp_path / "9.npz"), None, 1, 1500))
    msgs = _run(tasks, n_threads=2, batch=3)
    assert msgs[0][0] == "gpu" and msgs[-1] == ("exit", "gpu0")
    ok = {m[1][0]: m[1] for m in msgs if m[0] == "ok"}
    assert sorted(ok) == [0, 1, 2, 3, 4, 9]
    assert ok[9][1] == 0 and "FileNotFoundError" in ok[9][2]
    for idx, path, out, _, _, nfeat in tasks[:5]:
        _, n, err, tt = ok[idx]
        assert err is None and n > 10 and tt["photos_gpu"] == 1
        g = np.load(path)
        xy, desc, resp = per_image_path(g, nfeat)
        ref = str(tmp_path / ("ref%d.npz" % idx))
        assert features._save(ref, xy, desc, resp, g.shape[1], g.shape[0]) == n
        a, b = np.load(out), np.load(ref)
        for key in ("xy", "resp", "red_wh"):
            np.testing.assert_array_equal(a[key], b[key])
        assert np.abs(a["desc"].astype(int) - b["desc"].astype(int)).max() <= 1


@pytest.mark.gpu
def test_worker_gpu_clahe(tmp_path, gpu):
    """clahe=True: the GPU equalises the plain decoded photos; the feature files equal
    sift_gpu on the CPU restatement's CLAHE output."""
    from fieldstitch_amd import features
    from oracle import clahe_oracle
    shapes = [(240, 320), (300, 410), (240, 320)]
    tasks = _tasks(tmp_path, shapes)
    msgs = _run(tasks, n_threads=2, batch=2, clahe=True)
    ok = {m[1][0]: m[1] for m in msgs if m[0] == "ok"}
    assert sorted(ok) == [0, 1, 2]
    for idx, path, out, _, _, nfeat in tasks:
        _, n, err, _ = ok[idx]
        assert err is None and n > 10
        g = clahe_oracle.apply(np.load(path))
        xy, desc, resp = per_image_path(g, nfeat)
        ref = str(tmp_path / ("ref%d.npz" % idx))
        assert features._save(ref, xy, desc, resp, g.shape[1], g.shape[0]) == n
        a, b = np.load(out), np.load(ref)
        for key in ("xy", "resp", "red_wh"):
            np.testing.assert_array_equal(a[key], b[key])
        assert np.abs(a["desc"].astype(int) - b["desc"].astype(int)).max() <= 1


def _cpu_extract(task):
    """A stand-in for the reference's CPU extract (features.py:132-154): (idx, n, err, timings)."""
    return task[0], 1000 + task[0], None, {"photos_cpu": 1}


def test_worker_falls_back_to_cpu_without_gpu(tmp_path, monkeypatch):
    """No GPU and a cpu_fallback: every task of the reference's tuple form goes to the CPU
    path (the photo that cannot be decoded included), as features.py:355-364 does."""
    from fieldstitch_amd import features

    def no_gpu():
        raise RuntimeError("no ROCm GPU visible")
    monkeypatch.setattr(features, "_gpu_ready", no_gpu)
    tasks = _tasks(tmp_path, [(60, 80), (64, 96), (60, 80)])
    tasks.append((7, str(tmp_path / "missing.jpg"), str(tmp_path / "7.npz"), str(tmp_path / "t7.jpg"), 4, 1500))
    msgs = _run(tasks, cpu_fallback=_cpu_extract)
    assert msgs[0][0] == "gpu_off" and msgs[-1] == ("exit", "gpu0")
    ok = {m[1][0]: m[1] for m in msgs if m[0] == "ok"}
    assert sorted(ok) == [0, 1, 2, 7]
    assert all(r == (k, 1000 + k, None, {"photos_cpu": 1}) for k, r in ok.items())


def test_worker_gpu_failure_mid_run_continues_on_cpu(tmp_path, monkeypatch):
    """The GPU passes its check but fails on a batch: one "gpu_off" message, and that batch's
    photos and all later ones go to the CPU path."""
    from fieldstitch_amd import features
    monkeypatch.setattr(features, "_gpu_ready", lambda: "test device")

    class Broken:
        def __init__(self, *a, **k):
            raise RuntimeError("device lost")
    monkeypatch.setattr(features, "FeatureBatch", Broken)
    tasks = _tasks(tmp_path, [(60, 80), (64, 96), (60, 80), (64, 96), (60, 80)])
    msgs = _run(tasks, n_threads=1, batch=2, cpu_fallback=_cpu_extract)
    assert msgs[0][0] == "gpu"
    offs = [m for m in msgs if m[0] == "gpu_off"]
    assert len(offs) == 1 and "device lost" in offs[0][2]
    ok = {m[1][0]: m[1] for m in msgs if m[0] == "ok"}
    assert sorted(ok) == [0, 1, 2, 3, 4] and all(r[1] == 1000 + k for k, r in ok.items())


def test_worker_decode_task_gets_the_task_tuple(tmp_path, monkeypatch):
    from fieldstitch_amd import features

    def no_gpu():
        raise RuntimeError("no ROCm GPU visible")
    monkeypatch.setattr(features, "_gpu_ready", no_gpu)
    seen = []

    def dec(task):
        seen.append((task[0], task[3], task[4]))
        raise IOError("cannot decode image")
    tasks = _tasks(tmp_path, [(60, 80)])
    tasks[0] = tasks[0][:3] + ("thumb.jpg", 2, tasks[0][5])
    msgs = _run(tasks, decode_task=dec)
    ok = [m[1] for m in msgs if m[0] == "ok"]
    assert seen == [(0, "thumb.jpg", 2)] and ok[0][1] == 0 and "cannot decode image" in ok[0][2]


@pytest.mark.gpu
def test_worker_jpeg_flight_with_prepare(tmp_path, gpu):
    """JPEG photos through the worker as the pipeline feeds them: prepare_task decodes each
    reduced by the task's factor and writes its thumbnail, the GPU equalises (clahe=True) and
    extracts; the feature files equal sift_gpu on the CPU restatement's CLAHE of the same
    decoded image."""
    from PIL import Image
    from fieldstitch_amd import features, synth
    from oracle import clahe_oracle
    rng = np.random.default_rng(12)
    tasks = []
    for k, (h, w, r) in enumerate(((480, 640, 2), (600, 820, 2), (480, 640, 1))):
        mos = synth.Mosaic(seed=300 + k)
        mos.polys = synth.random_polygons(rng, 10, 0, 0, w, h)
        g = mos.render(np.eye(3)[:2], h, w, rng)
        rgb = np.stack([g, np.clip(g.astype(int) + 12, 0, 255), np.clip(g.astype(int) - 9, 0, 255)], -1)
        p = tmp_path / ("photo%d.jpg" % k)
        Image.fromarray(rgb.astype(np.uint8)).save(p, format="JPEG", quality=93)
        tasks.append((k, str(p), str(tmp_path / ("%d.npz" % k)), str(tmp_path / ("th%d.jpg" % k)), r, 1200))
    expect = {}
    for t in tasks:
        g = features.read_gray(t[1], t[4])
        expect[t[0]] = features.prepare_task((t[0], t[1], None, str(tmp_path / ("x%d.jpg" % t[0])), t[4], t[5]))
        assert expect[t[0]].shape == g.shape
    print("case jpeg worker", flush=True)
    msgs = _run(tasks, n_threads=2, batch=3, decode_task=features.prepare_task, clahe=True)
    ok = {m[1][0]: m[1] for m in msgs if m[0] == "ok"}
    assert msgs[0][0] == "gpu" and sorted(ok) == [0, 1, 2]
    for idx, path, out, thumb, r, nfeat in tasks:
        _, n, err, _ = ok[idx]
        assert err is None and n > 10
        with Image.open(thumb) as th:
            assert th.format == "JPEG"
        g = clahe_oracle.apply(expect[idx])
        xy, desc, resp = per_image_path(g, nfeat)
        ref = str(tmp_path / ("ref%d.npz" % idx))
        assert features._save(ref, xy, desc, resp, g.shape[1], g.shape[0]) == n
        a, b = np.load(out), np.load(ref)
        for key in ("xy", "resp", "red_wh"):
            np.testing.assert_array_equal(a[key], b[key])
        assert np.abs(a["desc"].astype(int) - b["desc"].astype(int)).max() <= 1
