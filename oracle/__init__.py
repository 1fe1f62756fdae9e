"""CPU FP64 oracle for the space-time chi0 -> RPA Ec and G0W0 Sigma path.

TEST INFRASTRUCTURE ONLY.  This package restates LibRPA's algorithm in numpy
so that the HIP implementation in ``librpa_amd`` can be checked against it.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / the timed CPU baseline;
nothing in the shipped library calls into it.

Every function cites the reference file:line it follows.  Pins (what fixes
each function independently of this restatement) are listed in the module
docstrings and exercised by ``tests/test_oracle_*.py``.

``Makefile`` builds three checkers from the reference source tree into ``_ref/``:
``ref_libri`` (``ref_libri.cpp``) runs the reference's own header-only LibRI contractions
(``RPA::cal_chi0s``, ``GW::cal_Sigmas``), ``ref_qpe`` (``ref_qpe.cpp``) its Pade continuation
and QP solver, ``ref_minimax`` (``ref_minimax.f90``) its vendored greenX minimax generator.
``tests/test_oracle_libri.py``, ``tests/test_oracle_ref_qpe.py`` and
``tests/test_oracle_ref_minimax.py`` hold this package against them.
"""
