"""CPU oracle for the frequency-domain beamforming path.

TEST INFRASTRUCTURE ONLY.  This package restates, in float64 numpy (``beamform.py``,
``spectra.py``) and in plain C (``cpu_ref.c``), the reference algorithms of Acoular's
frequency-domain beamforming path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or execute anything here, and only as the
checker or as the timed CPU baseline -- never as the thing measured or shipped.  The
product path in ``acoular_amd`` runs on the GPU and fails loudly when its HIP library is
missing; it never falls back to this code.

Every function cites the reference file:line it follows (paths relative to the Acoular
source tree).  The checks that pin each function (known-answer vectors, closed forms,
invariants, brute force) live in ``tests/test_oracle.py``.
"""
