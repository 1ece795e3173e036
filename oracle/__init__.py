"""CPU oracle of the ShowerTransport hot path — TEST INFRASTRUCTURE ONLY.

FP64 restatements of the reference algorithms (LightHit 0.3.0a9), each
function citing the reference file:line it follows. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker or the timed CPU baseline; the
``lighthit_amd`` product never imports, links or runs anything here.

Parity pins (what fixes each function independently of its own formula) are
listed in each module's docstring; a function without one says
"parity unpinned".
"""
