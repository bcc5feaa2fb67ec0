"""CPU oracle for the phasefield5d explicit step -- TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  It is used only by
``tests/`` (parity checks), ``__graft_entry__.smoke()`` (checking one device
step) and the ``cpu_baseline`` leg of ``bench.py`` (timing the CPU path).

* ``pf5d_oracle``  -- NumPy FP64 restatement of the hot-path functions,
  operation order kept per reference file:line so that FP64 results can be
  compared at 1e-12 (bit-exact where no FFT is involved).
* ``pf5d_oracle_c.c`` -- C/OpenMP FP64 restatement of the same loops, used as
  the CPU timing baseline (FFT through numpy.fft, as the reference uses
  scipy.fft/pocketfft).

Pinning (what checks each oracle function beyond retyping its formula) is
listed in the header of ``pf5d_oracle``.
"""
