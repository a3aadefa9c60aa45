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
* assembly / K.u     -- dense element loop with the geometry's own element nodes,
                        rigid-body modes, symmetry, energy identity, tip lump,
                        total mass (tests/test_oracle_fem.py)
* CG                 -- scipy.sparse.linalg.spsolve, true-residual bound;
                        Timoshenko tip deflection of a solid cantilever
* LOBPCG             -- dense scipy.linalg.eigh (eigenvalues, M-orthonormality,
                        MAC); the reference's committed ArmGeometry3D frequencies
                        (notes/solver_crosscheck.json:12), which are unconverged
                        upper bounds of tests/golden/arm_modal_exact.json
* sensitivities      -- central differences of exactly re-solved compliance and
                        eigenvalue (dense solves, both Pedersen branches, tip mass)
* filter             -- brute-force convolution loop, 19 taps and weight sum,
                        partition of unity
* OC / SIMP loop     -- the reference's own SIMPOptimizer run on a small
                        tip-loaded box (tests/golden/make_simp_golden.py,
                        tests/test_oracle_simp.py) and OC invariants (bounds,
                        move limit, volume, monotonicity)
* fatigue            -- the reference's committed OEM known-answer
                        (notes/oem_baseline.json:4,14,73-74) and hand cases
* effective mass     -- completeness (sum over all modes = 1) and a single rigid mode
"""
