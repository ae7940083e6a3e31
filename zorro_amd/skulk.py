"""GPU worker behind the Skulk daemon's host interface.

The reference's daemon (``automator/zorroSkulkManager.py``) hands each ready
movie to a ``skulkHost`` whose ``submitLocalJob`` writes the job's ``.zor``
file and starts ``zorro -c job.zor`` as a subprocess (``:1089-1113``); the
manager loop polls every host (``:1010-1062``, ``:1269-1370``) and reads the
state messages the host puts on its queue.  ``GpuSkulkHost`` keeps that
interface (constructor arguments, ``workerFunction(stateObj)``, ``poll()``
returning ``HOST_BUSY``/``HOST_FREE``, ``kill()``, the same messages) and runs
the job in-process instead: ``run_config`` is the CLI's sequence
(``zorro/__main__.py:49-103``: loadConfig + loadData, alignImageStack, saveData,
``METAstatus = 'fini'``, saveConfig) on the GPU, in a worker thread shared by
the hosts of one device, so no interpreter start-up, no re-import and no
CUDA-context creation is paid per movie and movies of consecutive jobs overlap
on the device.

Differences, deliberate: ``kill()`` cancels a job that has not started; a
running alignment cannot be interrupted and finishes, its result is dropped
(the reference kills the process tree).  PNG plots (``savePNG``) are not drawn.
"""
from __future__ import annotations

import os
import threading
import time
import traceback
from concurrent.futures import ThreadPoolExecutor

# state codes of automator/zorroSkulkManager.py:37-52; the manager compares them
NEW, CHANGING, STABLE, SYNCING, READY, PROCESSING, FINISHED, ARCHIVING, COMPLETE, STALE = range(10)
HOST_BUSY, HOST_FREE, HOST_ERROR, RENAME, ERROR, ZOMBIE = range(10, 16)
STATE_NAMES = {NEW: "new", CHANGING: "changing", STABLE: "stable", SYNCING: "sync", READY: "ready",
               PROCESSING: "proc", FINISHED: "fini", ARCHIVING: "archiving", COMPLETE: "complete",
               STALE: "stale", HOST_BUSY: "busy", HOST_FREE: "free", HOST_ERROR: "host_error",
               RENAME: "rename", ERROR: "error", ZOMBIE: "zombie"}


def run_config(config, stack=None, out_sum=None, registrator=None):
    """``zorro -c config [-i stack] [-o sum]`` in this process, on the GPU
    (``zorro/__main__.py:49-103``).  Returns the registrator."""
    if registrator is None:
        from .registrator import ImageRegistrator
        registrator = ImageRegistrator()
    reg = registrator
    if stack is None:
        reg.loadConfig(config, loadData=True)
    else:
        reg.loadConfig(config, loadData=False)
        reg.files["stack"] = stack
        reg.loadData()
    reg.bench["total0"] = time.time()
    if out_sum is not None:
        reg.files["sum"] = out_sum
    if reg.triMode == "refine":
        # 'refine' needs an alignment first unless the configuration holds one (:69-84)
        if not (reg.errorDictList and reg.errorDictList[-1]):
            reg.triMode = "diag"
            reg.alignImageStack()
            reg.loadData()
        reg.triMode = "refine"
    reg.alignImageStack()
    reg.saveData()
    reg.METAstatus = "fini"
    reg.saveConfig()
    return reg


_POOLS: dict = {}
_POOLS_LOCK = threading.Lock()


def _pool(device: int, workers: int) -> ThreadPoolExecutor:
    with _POOLS_LOCK:
        p = _POOLS.get(device)
        if p is None:
            p = _POOLS[device] = ThreadPoolExecutor(max_workers=workers, thread_name_prefix=f"zorro-gpu{device}")
        return p


class GpuSkulkHost:
    """``skulkHost`` (``automator/zorroSkulkManager.py:917-1203``) whose local
    jobs run on the GPU in-process.  ``workerFunction`` 'local' (and 'gpu')
    select the GPU job; 'dummy' sleeps as the reference's test job does; other
    reference workers (qsub, rsync, archive) are not built and raise."""

    def __init__(self, hostname, workerFunction, messageQueue, n_threads=None, cachepath=None, qsubHeader=None,
                 device=0, max_inflight=2, runner=run_config):
        self.hostName = hostname
        self.messageQueue = messageQueue
        self.zorroState = None
        self.submitName = None
        self.n_threads = n_threads
        self.cachePath = cachepath
        self.qsubHeaderFile = qsubHeader
        self.device = device
        self._max_inflight = max_inflight
        self._runner = runner
        self._future = None
        self.lastError = None          # exception of the last failed job
        if workerFunction in ("local", "gpu"):
            self.workerFunction = self.submitLocalJob
        elif workerFunction == "dummy":
            self.workerFunction = self.submitDummyJob
        elif callable(workerFunction):
            self.workerFunction = workerFunction
        else:
            raise NotImplementedError(f"skulk worker {workerFunction!r}: only local GPU jobs are built")

    def _clean(self):
        self._future = None
        self.zorroState = None

    def submitLocalJob(self, stateObj):
        """Write the job's configuration (METAstatus 'proc', stdout log beside
        it) and queue the GPU job; puts ``[HOST_BUSY, stateObj]``."""
        self.zorroState = stateObj
        z = stateObj.zorroObj
        z.n_threads = self.n_threads
        if self.cachePath is not None:
            z.cachePath = self.cachePath
        z.files["stdout"] = os.path.splitext(z.files["config"])[0] + ".zout"
        z.METAstatus = STATE_NAMES[PROCESSING]
        z.saveConfig()
        cfg, log = z.files["config"], z.files["stdout"]
        runner = self._runner

        def job():
            # the subprocess host appends the job's output to files['stdout']
            # (zorro -c cfg >> stdout 2>&1); a failing job leaves its traceback there
            try:
                return runner(cfg)
            except BaseException:
                with open(log, "a") as fh:
                    fh.write(traceback.format_exc())
                raise

        self.lastError = None
        self._future = _pool(self.device, self._max_inflight).submit(job)
        self.messageQueue.put([HOST_BUSY, self.zorroState])

    def submitDummyJob(self, stateObj):
        self.zorroState = stateObj
        time.sleep(5)

    def poll(self):
        """``HOST_BUSY`` while the job runs; on success the state is advanced
        (``incrementState``) and announced, on failure ``[HOST_ERROR, state]``
        is sent; both then free the host (``:1010-1062``)."""
        if self.zorroState is None:
            return HOST_FREE
        if self._future is None:
            self._clean()
            return HOST_FREE
        if not self._future.done():
            return HOST_BUSY
        st = self.zorroState
        failed = self._future.cancelled() or self._future.exception() is not None
        if failed:
            self.lastError = None if self._future.cancelled() else self._future.exception()
            self.messageQueue.put([HOST_ERROR, st])
        else:
            st.incrementState()
            self.messageQueue.put([st.state, st])
        self._clean()
        return HOST_FREE

    def kill(self):
        """Cancel a queued job; a running one finishes and its result is dropped."""
        if self._future is not None:
            self._future.cancel()
        self._clean()
