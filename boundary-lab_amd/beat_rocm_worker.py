"""Executable worker script: drop-in ``solver_script`` for Boundary Lab's BeatEngineBackend.

Usage from Boundary Lab:
    BeatEngineRocmBackend(julia_executable=sys.executable, solver_script="<repo>/boundary-lab_amd/beat_rocm_worker.py",
                          julia_project=None)
which runs ``python beat_rocm_worker.py --worker`` and speaks the reference worker's NDJSON protocol.
"""

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from beat_hip.worker import main  # noqa: E402

if __name__ == "__main__":
    sys.exit(main())
