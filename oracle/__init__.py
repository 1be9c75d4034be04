"""CPU oracle for the MOAO reconstructor path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference's algorithm for the hot path so the HIP
product can be checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker / the timed CPU
baseline.  The product (``moao_amd``) never imports, links or calls anything here.

Modules
  params.py      reference text formats (sys-params / atm-params)      common/io/*, common/tomo_struct/{sys,atm}Params.cpp
  tomo.py        Tomo_struct preparation (subapertures, projection, noise)  common/tomo_struct/tomo_struct.cpp
  matcov_oracle.c + matcov.py   matcov element / tile generator       common/kernels/matcov_kernels.cpp, common/compute/matcov.cpp
  lapack_path.py reconstructor and Cee_Cvv in the reference's LAPACK/BLAS call order
                                                                       lapack/compute/{reconstructor,cee_cvv}.cpp
  intersample.py PSF from Cvv with numpy FFTs                          common/tomo_struct/intersample.cpp
  check.py       compareMatrices and relative Frobenius error          common/compute/check_utils.cpp

How each function is pinned (golden vectors from the hand-built reference, closed forms,
library routines) is listed in the header of each module.
"""
