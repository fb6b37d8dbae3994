"""FP64 CPU oracle for gopf's spectral time step.

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
it.  The product library (``gopf_amd``) never imports, links or calls it.

Parity pins (each function's independent check lives in tests/test_oracle_*.py):
  * FFT (fftWrap.go:16-39)            brute-force DFT, numpy.fft, delta / single-mode
  * Freq / ConjugateNode (42-95)      closed form + TestConjugateNode restated
  * parser (rhsBuilder.go, util.go)   the reference's parser tests restated
  * Euler / RK4 / Solver              TestExponentialDecay, TestEulerSquareDecary,
                                      TestSimpleModel, TestWithImplicit,
                                      TestSolverDiffusion restated (closed forms)
  * SquaredGradient                   TestSquareGradient / TestSquareGradWithSolver
  * HomogeneousModulusLinElast        TestHomogeneousRHS (closed form with injected
                                      displacements) + isotropic closed form K(k)
  * Vandeven / ApplyModalFilter       TestVandeven, TestApplyModalFilter, incomplete beta
  * CH energy (no reference)          single-mode closed form
"""
