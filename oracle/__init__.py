"""CPU oracle for the DD projector pair and FDK filter.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; never by the
product package ``ctreconstruction_amd``.
"""
