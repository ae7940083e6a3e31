"""FP64 CPU oracle for the Zorro drift-registration path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (``zorro_amd``) may
import, call or execute anything in this directory. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or the timed CPU baseline), never as the measured or
shipped path.

The oracle restates, in float64 NumPy/SciPy, the algorithm of
``zorro/zorro.py`` (``ImageRegistrator.alignImageStack`` with
``xcorrMode='zorro'``) and the helpers it calls in ``zorro/zorro_util.py``.
Every function cites the reference lines it follows.

Pinning status per function is listed in ``zorro_oracle.PINS`` and repeated
in the tests that pin them (``tests/test_oracle_pins.py``).
"""
from .zorro_oracle import *  # noqa: F401,F403
