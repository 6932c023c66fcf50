"""CPU restatement of the AMFLS block Golub-Kahan least-squares path.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP engine in
``amfls_solver_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to compute the
expected answer (or to time the CPU baseline).  The product never calls it.

It restates, in numpy/scipy FP64 (plus numpy ``longdouble`` for the outward
rounded bound arithmetic), the reference's algorithm:

* ``fpbounds``   -- src/algorithms/mathematics/floating_point_bounds.hpp
* ``rng``        -- src/algorithms/mathematics/random/gaussian.cpp
* ``schedule``   -- src/algorithms/mathematics/krylov/certificate_checkpoint_schedule.hpp
* ``controller`` -- src/algorithms/amfls/amfls_internal.hpp
* ``operators``  -- src/algorithms/mathematics/operators/*.cpp
* ``validation`` -- src/algorithms/mathematics/krylov/candidate_validation.cpp
* ``workspace``  -- src/algorithms/mathematics/krylov/krylov_workspace.cpp
* ``solvers``    -- src/algorithms/amfls/amfls.cpp, src/algorithms/lsqr/lsqr.cpp,
                    src/algorithms/fixed_rbgk/fixed_rbgk.cpp
* ``results``    -- include/amfls/least_squares_result.hpp,
                    bindings/torch/result_tensors.hpp

LAPACK semantics: numpy.linalg.svd(full_matrices=False) is LAPACK dgesdd
(the reference's thin_svd) and scipy.linalg.lstsq(lapack_driver="gelsy") is
LAPACK dgelsy (the reference's minimum_norm_least_squares).

Pinning status of each function is recorded in its module docstring; the
checks live in tests/test_oracle_*.py.
"""
