"""CPU oracle for the implicit-FEM backward-Euler path — TEST INFRASTRUCTURE ONLY.

This package is a plain numpy/scipy FP64 restatement of the reference's
algorithm (simkit v0.1.4). Each function cites the reference file:line it
follows. It exists to check the HIP library, never to stand in for it:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or the timed CPU
baseline). The product package ``simkit_amd`` never imports it.

Pinning status of each function (what independently fixes its behaviour)
is listed in the module docstrings and enforced by ``tests/test_oracle*.py``:

* ``mesh``      — counts, volume sum, vertex ordering, golden fixture.
* ``fem``       — F/J: affine-map exact recovery; SNH: independent sympy
                  derivation, finite differences, golden fixture from the
                  reference; ARAP/SVD: polar invariants, FD, golden fixture;
                  psd_project: eigen invariants, golden fixture; assembly:
                  brute-force dense assembly, golden CSR fixture.
* ``solver``    — line search / Newton: the reference's known-answer tests;
                  BE: SHO closed form, golden trajectory fixture.
"""
