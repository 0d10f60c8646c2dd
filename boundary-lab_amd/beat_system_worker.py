"""Executable worker script: drop-in ``solver_script`` for Boundary Lab's physical-system backend on ROCm.

Usage from Boundary Lab:
    PhysicalSystemProductionBackend(bem_backend="rocm", julia_executable=sys.executable,
                                    solver_script="<repo>/boundary-lab_amd/beat_system_worker.py")
which runs ``python beat_system_worker.py --worker`` and speaks the coupled worker's NDJSON protocol for
exterior (unbounded-region-only) systems.
"""

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from beat_hip.system_worker import main  # noqa: E402

if __name__ == "__main__":
    sys.exit(main())
