"""FP64 CPU oracle for the jVMC VMC inner loop -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy (float64 / complex128), the algorithms of
jVMC 1.5.8 that the HIP library replaces.  Each function cites the reference
file:line it follows.  It is the checker against which the GPU path is
compared, and the CPU baseline timed beside it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it;
the product package ``vmc_jax_amd`` never does.

Parity pinning (what fixes each part independently of this restatement):

* ``philox``      Random123 known-answer vectors (tests/test_oracle_philox.py).
* ``nets``        closed forms, CNN translation invariance (nets_test.py:16-46),
                  central finite differences for every gradient leaf.
* ``operator``    dense Hamiltonians built from numpy ``kron`` (brute force),
                  known answers of operator_test.py:48-93.
* ``sampler``     exact-sampler closed form (sampler.py:442-486); MCMC histogram
                  against exact probabilities (sampler_test.py:32-82 method).
* ``stats``       closed-form values of stats_test.py:15-38.
* ``solvers``     ED ground-state energies (tdvp_test.py:32, minsr_test.py:20),
                  the ZZ(t) golden curve (tdvp_test.py:122-126), AdaptiveHeun
                  against ``scipy.linalg.expm`` (stepper_test.py:14-39), the SNR
                  identity (tdvp_test.py:201-263).

Parity unpinned: the JAX threefry bit streams (replaced by Philox4x32-10 here and
on the device), Flax initial parameter values, last-bit results of XLA
reductions, and the eigenvector basis inside degenerate eigenspaces.
"""
