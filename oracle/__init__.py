"""CPU oracle for the CA-QR hot path (test infrastructure only).

Nothing in the product path (``candmc_amd``) imports this package.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only to check results against or as the timed
CPU baseline.
"""
