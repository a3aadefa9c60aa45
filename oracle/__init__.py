"""FP64 CPU oracle for the voxel hex8 FEM hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``vox8``) imports,
calls or executes anything in this directory.  It is used by

* ``tests/`` as the parity checker,
* ``__graft_entry__.smoke()`` as the checker of one small GPU invocation,
* ``bench.py``'s ``cpu_baseline`` leg, timed beside the GPU numbers.

It is a NumPy/SciPy restatement, written for this project, of the semantics of
the reference's ``src/fem.py``, ``src/topopt.py``, ``src/fatigue.py``,
``src/mesh3d.py`` and ``sims/modal_participation.py``.  Every function cites the
reference ``file:line`` it follows.  No reference file is imported.

Pinning (how each part is checked beyond its own formula):

* element matrices   -- exact rational integration (sympy), rigid-body null space,
                        mass row sums (tests/test_oracle_element.py)
* assembly / K.u     -- brute-force element loop on tiny grids, symmetry,
                        energy identity (tests/test_oracle_fem.py)
* CG                 -- scipy.sparse.linalg.spsolve, residual bound
* LOBPCG             -- dense scipy.linalg.eigh on tiny grids; the reference's
                        committed ArmGeometry3D frequencies
                        (notes/solver_crosscheck.json:12)
* sensitivities      -- central finite differences of stationary functionals
* filter             -- brute-force convolution loop, interior weight sum
* OC / SIMP loop     -- invariants (volume, bounds, move limit) and golden
                        vectors written by tests/golden/make_simp_golden.py
* fatigue            -- the reference's committed OEM known-answer
                        (notes/oem_baseline.json:4,14,73-74) and hand cases
* effective mass     -- closed form on rigid translation
"""
