"""CPU oracle of the FDTD step.  TEST INFRASTRUCTURE: imported only by tests/, by
``__graft_entry__.smoke()`` (as the checker) and by the ``cpu_baseline`` leg of bench.py."""
