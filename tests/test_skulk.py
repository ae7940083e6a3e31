"""The GPU worker behind the Skulk host interface (automator/zorroSkulkManager.py:
917-1203): state messages, poll results and error handling, checked on the CPU
with a stand-in job; the end-to-end GPU job (MRC movie -> .zor -> sums on disk)
is compared with a direct ImageRegistrator run on the same file."""
import os
import queue
import time

import numpy as np
import pytest

from zorro_amd import skulk
from zorro_amd.registrator import ImageRegistrator


class _State:
    """The parts of zorroState (zorroSkulkManager.py:69-161) a host uses."""

    def __init__(self, zobj, q):
        self.zorroObj = zobj
        self.messageQueue = q
        self.state = skulk.PROCESSING

    def incrementState(self):
        if self.state == skulk.PROCESSING:
            self.state = skulk.FINISHED
            self.messageQueue.put([self.state, self])


def _wait_free(host, limit=120.0):
    t0 = time.time()
    while host.poll() == skulk.HOST_BUSY:
        assert time.time() - t0 < limit, "job did not finish"
        time.sleep(0.05)


def test_state_codes_match_the_manager():
    assert (skulk.PROCESSING, skulk.FINISHED, skulk.HOST_BUSY, skulk.HOST_FREE, skulk.HOST_ERROR) == (5, 6, 10, 11, 12)
    assert skulk.STATE_NAMES[skulk.FINISHED] == "fini" and skulk.STATE_NAMES[skulk.PROCESSING] == "proc"


def test_host_messages_success_and_failure(tmp_path):
    ran = []

    def job(cfg):
        ran.append(cfg)
        time.sleep(0.2)
        if "bad" in cfg:
            raise RuntimeError("alignment failed")

    q = queue.Queue()
    host = skulk.GpuSkulkHost("gpu0", "local", q, n_threads=4, runner=job, device=99)
    assert host.poll() == skulk.HOST_FREE
    z = ImageRegistrator()
    z.files["stack"] = str(tmp_path / "good.mrcs")
    z.files["config"] = str(tmp_path / "good.zor")
    st = _State(z, q)
    host.workerFunction(st)
    assert q.get_nowait() == [skulk.HOST_BUSY, st]
    assert host.poll() == skulk.HOST_BUSY
    # the job's configuration is on disk with METAstatus as its second line (zorroState.peek)
    with open(z.files["config"]) as fh:
        fh.readline()
        assert fh.readline().split() == ["METAstatus", "=", "proc"]
    assert z.files["stdout"] == str(tmp_path / "good.zout") and z.n_threads == 4
    _wait_free(host)
    assert q.get_nowait() == [skulk.FINISHED, st] and q.get_nowait() == [skulk.FINISHED, st]
    assert ran == [z.files["config"]]
    # a failing job reports HOST_ERROR and frees the host
    zb = ImageRegistrator()
    zb.files["config"] = str(tmp_path / "bad.zor")
    sb = _State(zb, q)
    host.workerFunction(sb)
    q.get_nowait()
    _wait_free(host)
    assert q.get_nowait() == [skulk.HOST_ERROR, sb] and sb.state == skulk.PROCESSING
    assert isinstance(host.lastError, RuntimeError)
    assert "alignment failed" in open(tmp_path / "bad.zout").read()
    assert host.poll() == skulk.HOST_FREE and q.empty()
    with pytest.raises(NotImplementedError):
        skulk.GpuSkulkHost("h", "qsub", q)


@pytest.mark.gpu
def test_gpu_job_matches_direct_run(tmp_path, monkeypatch):
    from zorro_amd import mrc
    from zorro_amd.synth import make_movie
    monkeypatch.chdir(tmp_path)
    n, h, w = 8, 1024, 1024
    print(f"[gpu skulk] job n={n} {h}x{w}", flush=True)
    frames, _ = make_movie(n, h, w, seed=71, dose=1.0)
    mrc.write_mrc("m1.mrcs", frames.astype(np.uint16), pixelsize_nm=0.1)
    settings = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, fouCrop=[1024, 1024], shapePadded=[1024, 1024],
                    filterMode="dose", saveMovie=False)
    z = ImageRegistrator()
    for k, v in settings.items():
        setattr(z, k, v)
    z.files["stack"] = "m1.mrcs"
    z.files["config"] = "m1.zor"
    q = queue.Queue()
    host = skulk.GpuSkulkHost("gpu0", "local", q)
    st = _State(z, q)
    host.workerFunction(st)
    _wait_free(host, limit=300.0)
    msgs = [q.get_nowait() for _ in range(q.qsize())]
    if os.path.exists("m1.zout"):
        print(open("m1.zout").read(), flush=True)
    assert [m[0] for m in msgs] == [skulk.HOST_BUSY, skulk.FINISHED, skulk.FINISHED], repr(host.lastError)
    done = ImageRegistrator().loadConfig("m1.zor")
    assert done.METAstatus == "fini" and done.files["sum"] == os.path.join("sum", "m1_zorro.mrc")
    direct = ImageRegistrator()
    for k, v in settings.items():
        setattr(direct, k, v)
    direct.loadData("m1.mrcs")
    direct.alignImageStack()
    assert np.array_equal(done.translations, direct.translations)
    s, hdr = mrc.read_mrc(done.files["sum"])
    assert np.array_equal(s[0], direct.imageSum) and hdr["pixelsize"][-1] == pytest.approx(0.1, rel=1e-6)
    f, _ = mrc.read_mrc(done.files["filt"])
    assert np.array_equal(f[0], direct.filtSum)
    assert np.array_equal(done.FRC, direct.FRC) and done.FRC.size > 0
    direct.close()
