"""CPU restatement of the tamc sampling path (test infrastructure only).

Imported by tests/, the smoke check and bench.py's cpu_baseline leg, never by the
product package.  See tamc_oracle.cpp for the reference file:line each function restates
and for the pin that checks it.
"""
from .oracle import *  # noqa: F401,F403
