"""CPU oracle of Ariel's FLIP step — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
and there only as the checker. The product (ariel_amd) never imports it.
"""
from .oracle import OracleSim, load, build  # noqa: F401
