"""FP64 CPU oracle for the inverse-lithography hot path.

TEST INFRASTRUCTURE ONLY.  This package is a numpy restatement of the
reference algorithm (wyhcoder/C-version-Lithography, "Litho_cpp") used as the
parity checker for the HIP path and as the timed CPU baseline.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product library never calls it.

Modules (each function cites the reference file:line it restates):

* ``optics``    Grid / Pupil / Zernike / Source / SOCS kernel precompute
                (grid.cpp, pupil.cpp, zernike.cpp, source.cpp, litho_prepare.cpp)
* ``imaging``   SOCS forward, JVP, PE losses and adjoint gradients
                (imaging.cpp, gradient.cpp, loss.cpp)
* ``ctm``       CTM theta init, gradient descent, L-BFGS (CTM_Optimizer.cpp)
* ``levelset``  WENO5-JS / Godunov / curvature / CFL / chamfer SDF / renderer
                (level_set_utils.cpp, levelset_mask_renderer.cpp)
* ``fullchip``  Partition, Micro gradient exchange, full-chip CTM and LSM
                (partition.cpp, micro_gradient_exchange.cpp, full_chip_ctm.cpp,
                full_chip_lsm.cpp)

Parity pins (what each part is checked against, in tests/):

* FFT conventions: numpy/pocketfft FP64 and brute-force DFT at tiny N.
* optics: closed forms (axis values, window 55 @ N=512 and 219 @ N=2048),
  SVD energy identity, SOCS -> Abbe convergence as K -> Ns.
* imaging: known answers (M=0, M=1, single pixel), shift equivariance,
  brute-force DFT at N <= 16, adjoint dot-product test, finite differences,
  JVP cross-check.
* ctm: the reference optimizer-control invariants (test_ctm_optimizer.cpp).
  L-BFGS line-search trajectory: PARITY UNPINNED (LBFGS++ is not vendored).
* levelset: the reference Godunov tests (test_level_set_godunov.cpp).
  Chamfer SDF re-init: PARITY UNPINNED against OpenCV (absent); the
  restatement is pinned by exact shortest paths on the 5x5 chamfer graph.
  Gray rendering: PARITY UNPINNED against cv::resize (absent); the
  restatement is pinned by closed forms (scale 1, constant and linear fields).
* fullchip: the reference exchange / trajectory tests
  (test_micro_gradient_exchange.cpp, test_full_chip_ctm.cpp,
  test_full_chip_lsm.cpp, test_mask_continuity.cpp).
"""
