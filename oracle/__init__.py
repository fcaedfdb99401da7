"""CPU oracle for phaser's reconstruction hot path -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch numpy restatement of the numpy code path of phaser
(wdwzyyg/phaser): FFT conventions, object sampling and cutouts, optics, groupings,
noise models, the LSQ-ML / ePIE group steps and the conventional engine loop, and the
forward model and loss of the gradient engine.  Every function cites the reference
file:line it follows.

It exists only to check the HIP backend (`phaser_amd`).  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it, and
only as the checker or the timed CPU baseline, never as part of the product path.

Two precisions are available from the same code: pass `dtype=numpy.float32` for the
FP32 mirror (same ops and dtype promotions as the reference numpy backend) and
`dtype=numpy.float64` for the FP64 twin used to state errors against FP64.

Parity pins (see tests/test_oracle_*.py):
  * misc.create_sparse_groupings / create_compact_groupings / create_rng: reference
    known answers tests/test_misc.py:7-49 and tests/test_object.py:12-29.
  * num.fft2 / ifft2: closed forms tests/test_num.py:102-155, naive DFT brute force,
    Parseval.
  * object.ObjectSampling: known answers tests/test_object.py:31-144, half-to-even ties
    by brute force; cutouts: ramp corners tests/test_object.py:266-296, sum invariant
    tests/test_object.py:243-244, brute-force slicing.
  * optics: propagator sign invariant tests/test_optics.py:44-57, unit-norm probe,
    propagate(+dz) then propagate(-dz) identity inside the band limit.
  * physics.Electron: closed forms tests/test_physics.py:7-49.
  * solvers (lsqml_run, epie_run, dry runs) and engine.run_engine: no reference test
    exercises them (parity unpinned by the reference).  Pinned by: FP32 vs FP64 twin,
    the data==model fixed point, S=1 reduction, a brute-force explicit-loop step on tiny
    inputs, and closed-form alpha step sizes.
"""
