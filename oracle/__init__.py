"""CPU restatement of the reference engine.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product path
(``analyticmesh_amd``) never imports it.
"""
from .runner import OracleRun, run_oracle, load_oracle, ensure_built  # noqa: F401
