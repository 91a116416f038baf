"""CPU oracle for the DistMesh force loop (test infrastructure only).

This package is a plain numpy FP64 restatement of the reference's algorithm
(SeismicMesh 3.6.2).  It is the checker that the HIP path in
``seismicmesh_amd`` is compared against, and the CPU baseline that the bench
times.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import it; the product never does.

Modules and what pins them (pins live in ``tests/test_oracle_*.py``):

* ``oracle.geometry`` -- SDF primitives and CSG
  (signed_distance_functions.py:70-623), box SDFs (fast_geometry.cpp:104-185),
  dihedral angles (fast_geometry.cpp:351-418), circumsphere gradient
  (fast_geometry.cpp:580-666), simp_vol / simp_qual / fix_mesh and boundary
  helpers (geometry/utils.py:175-469).  Pinned against closed forms, the
  reference's known answers (tests/test_geometry.py:44-56) and the
  reference's own compiled ``_fast_geometry`` module in ``oracle/_ref``.
* ``oracle.sizing`` -- gridded sizing (mesh_size_function.py:391-523) and
  gradient limiting (FastHJ.cpp:63-157).  Pinned against scipy's
  RegularGridInterpolator, the reference tests' interpolant values
  (tests/test_2dmesher_vs_water.py:42-43) and the compiled ``_FastHJ``.
* ``oracle.distmesh`` -- the generate_mesh / sliver_removal loops
  (mesh_generator.py:59-915, generation/utils.py:15-68).  Triangulation uses
  scipy's Qhull as a stand-in for CGAL (absent here); connectivity is pinned
  by an exact empty-circumcircle brute force (``oracle.predicates``).
* ``oracle.predicates`` -- exact orientation / incircle tests in rational
  arithmetic (Python ``fractions``), for brute-force Delaunay checks on small
  inputs.

Parity unpinned (no golden vector or closed form reaches them): CGAL's vertex
enumeration order and its handling of degenerate (cocircular) inputs, and
pyamg's approximate Laplacian solve used only in termination.
"""
