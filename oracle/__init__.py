"""CPU oracle for the dense-operator Krylov path — TEST INFRASTRUCTURE ONLY.

NumPy/SciPy FP64 (and FP32 where the reference computes in FP32)
restatements of the reference algorithms, each citing the reference file:line
it follows. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import this package, and only as the checker or the timed CPU
baseline; the product (mpdok_amd) never calls it.

Pinning (what checks each function beyond retyping its formula) is listed
function by function in tests/test_oracle.py:
  rbf.*          brute-force distances, closed-form kernel values, diag = k(0)+reg,
                 symmetry, the uniform-cube known answer of E[d^2]
  krylov.*       textbook special cases (A = I, k distinct eigenvalues -> breakdown
                 at step k, b = 0), history-shape rules, agreement with SciPy
                 cg/gmres/LAPACK solutions
  lu.*           P A = L U reconstruction, scipy.linalg.lu_factor on tiny N,
                 solve_triangular, LU-IR convergence on well-conditioned systems
  ozaki.*        digits reconstruct the rounded input exactly (integer sums),
                 bit-exact integer products, error bound 2^(-(8 s - 1)) scaling,
                 digit ranges at the int8 limits
  residual.*     exact rational arithmetic (fractions.Fraction) on small inputs
"""
