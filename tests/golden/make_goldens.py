"""Writes the committed fixtures under tests/golden/ (run in a container that
has the reference checkout at /root/reference and scipy).

* refs_*.npz   -- copies of the reference's own golden data files
                  (tests/refs/<name>.npz: end states Q, Qaux of whole marches).
* rcm_*.npz    -- CSR graphs of small triangle meshes with the permutation of
                  scipy.sparse.csgraph.reverse_cuthill_mckee.
"""
import pathlib
import shutil
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
REFS = pathlib.Path("/root/reference/tests/refs")
NAMES = ["parity", "stoker_wet", "ritter_dry", "wb_lake", "bc_wall",
         "swe_2d", "stoker_wet_o2_small", "ritter_dry_o2_small", "bc_periodic",
         "bc_inflow", "boundary_o2_small", "swashes_stoker_wet_o1",
         "swashes_stoker_wet_o2", "swashes_ritter_dry_o1", "boundary_order2",
         "wb_long"]


def copy_refs():
    for n in NAMES:
        shutil.copyfile(REFS / f"{n}.npz", HERE / f"refs_{n}.npz")


def rcm_fixtures():
    sys.path.insert(0, str(HERE.parents[1] / "zoomy-jax_amd"))
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    from zoomy_hip.mesh import delaunay_mesh, jittered_points
    for seed, (nx, ny) in ((1, (9, 7)), (2, (13, 11)), (3, (21, 13))):
        pts = jittered_points(nx, ny, (0.0, 10.0, 0.0, 5.0), 0.3, seed)
        m = delaunay_mesh(pts, (0.0, 10.0, 0.0, 5.0))
        indptr, indices = m.adjacency_csr()
        N = m.n_cells
        A = csr_matrix((np.ones(len(indices)), indices, indptr), shape=(N, N))
        # scipy seeds each component with np.argsort(degree)[0], an unstable
        # sort.  Make the minimum degree unique (extra edges from the other
        # minimum-degree nodes to a node two hops away) so that the fixture
        # pins the algorithm, not numpy's tie-break inside its sort.
        A = A.tolil()
        while True:
            deg = np.array([len(r) for r in A.rows])
            mins = np.nonzero(deg == deg.min())[0]
            if len(mins) == 1:
                break
            for v in mins[1:]:
                nb = A.rows[v][0]
                two = [w for w in A.rows[nb] if w != v and w not in A.rows[v]]
                w = two[0]
                A[v, w] = 1.0
                A[w, v] = 1.0
        A = A.tocsr()
        A.sort_indices()
        indptr, indices = A.indptr.astype(np.int32), A.indices.astype(np.int32)
        perm = reverse_cuthill_mckee(A, symmetric_mode=True)
        np.savez_compressed(HERE / f"rcm_seed{seed}.npz", indptr=indptr,
                            indices=indices, perm=perm.astype(np.int32),
                            points=pts)


def rcm_tied_fixture():
    """scipy's RCM of the adjacency of a reduced wet/dry triangle mesh
    (Delaunay of 100 x 50 jittered points on [0, 10] x [0, 5], seed 3), ties
    of the degree left in place: scipy starts from np.argsort(degree)[0],
    whose choice among equal degrees is numpy's unstable sort; the fixture
    records the result and its start node perm[-1]."""
    sys.path.insert(0, str(HERE.parents[1] / "zoomy-jax_amd"))
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    from zoomy_hip.mesh import delaunay_mesh, jittered_points
    pts = jittered_points(100, 50, (0.0, 10.0, 0.0, 5.0), 0.3, 3)
    m = delaunay_mesh(pts, (0.0, 10.0, 0.0, 5.0))
    indptr, indices = m.adjacency_csr()
    N = m.n_cells
    A = csr_matrix((np.ones(len(indices)), indices, indptr), shape=(N, N))
    A.sort_indices()
    perm = reverse_cuthill_mckee(A, symmetric_mode=True)
    np.savez_compressed(HERE / "rcm_tied_wetdry.npz",
                        indptr=A.indptr.astype(np.int32),
                        indices=A.indices.astype(np.int32),
                        perm=perm.astype(np.int32))


if __name__ == "__main__":
    copy_refs()
    rcm_fixtures()
    rcm_tied_fixture()
    print("fixtures written to", HERE)
