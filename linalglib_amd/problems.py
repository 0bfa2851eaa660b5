"""Synthetic SPD test problems in CSR form (int32 indices, float64 values).

The reference ships no generators (its tests read downloaded SuiteSparse
files), so these are this package's own.  Every generator returns
``(row_ptr, col_ind, val, m)`` with rows ordered lexicographically
(x fastest) and columns ascending within each row.

* ``poisson2d`` / ``poisson3d``: 5- and 7-point Laplacians with Dirichlet
  boundaries by truncation (diagonal 4 / 6, off-diagonals -1).
* ``diffusion3d``: 7-point finite-volume diffusion with per-face
  conductances; ``cell_kappa`` gives harmonic face averaging
  ``2 k_p k_q / (k_p + k_q)`` and Dirichlet faces use ``k_p``.
* ``convection_diffusion2d``: the nonsymmetric 5-point operator
  ``-lap u + w . grad u`` on the unit square (mesh width 1/(n+1), first-order
  upwinding of the convection term): an M-matrix, for the nonsymmetric
  Krylov solvers.
"""
import numpy as np


def poisson2d(nx, ny=None):
    ny = nx if ny is None else ny
    return _grid_laplacian((nx, ny, 1), dim=2)


def poisson3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _grid_laplacian((nx, ny, nz), dim=3)


def _neighbours(shape, dim):
    """Ordered (axis, step) list so that column offsets are ascending."""
    nx, ny, nz = shape
    out = []
    if dim == 3:
        out.append((2, -1))
    out.append((1, -1))
    out.append((0, -1))
    out.append(None)  # diagonal
    out.append((0, 1))
    out.append((1, 1))
    if dim == 3:
        out.append((2, 1))
    return out


def _grid_laplacian(shape, dim, face_k=None, diag_extra=None):
    """Assemble a 5/7-point operator.

    ``face_k`` maps (axis, step) -> conductance array of grid shape (k of the
    face between a cell and its neighbour in that direction; for a missing
    neighbour it is the Dirichlet face conductance).  None means unit
    conductances with a Dirichlet diagonal equal to the full stencil weight,
    i.e. the constant-coefficient Laplacian.
    """
    nx, ny, nz = shape
    m = nx * ny * nz
    idx = np.arange(m, dtype=np.int64).reshape(nz, ny, nx)
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    coords = {0: i3, 1: j3, 2: k3}
    sizes = {0: nx, 1: ny, 2: nz}
    strides = {0: 1, 1: nx, 2: nx * ny}

    nbrs = _neighbours(shape, dim)
    cols = []
    vals = []
    present = []
    diag = np.zeros((nz, ny, nx), dtype=np.float64)
    for nb in nbrs:
        if nb is None:
            continue
        ax, st = nb
        c = coords[ax]
        ok = (c + st >= 0) & (c + st < sizes[ax])
        kf = np.ones((nz, ny, nx)) if face_k is None else face_k[(ax, st)]
        diag += kf  # every face (interior or Dirichlet) adds to the diagonal
        cols.append(idx + st * strides[ax])
        vals.append(-kf)
        present.append(ok)
    if face_k is None:
        diag[...] = 2.0 * dim
    if diag_extra is not None:
        diag += diag_extra

    # interleave in ascending column order with the diagonal in the middle
    order = []
    half = len(cols) // 2
    for t in range(half):
        order.append((cols[t], vals[t], present[t]))
    order.append((idx, diag, np.ones_like(present[0])))
    for t in range(half, len(cols)):
        order.append((cols[t], vals[t], present[t]))

    C = np.stack([o[0].reshape(-1) for o in order], axis=1)
    V = np.stack([o[1].reshape(-1) for o in order], axis=1)
    P = np.stack([o[2].reshape(-1) for o in order], axis=1)
    counts = P.sum(axis=1)
    row_ptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    col_ind = C[P].astype(np.int32)
    val = V[P].astype(np.float64)
    return row_ptr.astype(np.int32), col_ind, val, m


def diffusion3d(nx, cell_kappa=None, face_kappa=None):
    """Variable-coefficient 7-point FV diffusion on an nx^3 grid.

    Exactly one of ``cell_kappa`` (grid-shaped cell conductivities; faces get
    harmonic means, Dirichlet faces the cell value) or ``face_kappa`` (dict
    (axis, step) -> grid-shaped face conductance) must be given.
    """
    shape = (nx, nx, nx)
    if face_kappa is None:
        k = np.asarray(cell_kappa, dtype=np.float64).reshape(nx, nx, nx)
        face_kappa = {}
        for ax_np, ax in ((2, 0), (1, 1), (0, 2)):
            for st in (-1, 1):
                nb = np.roll(k, -st, axis=ax_np)
                h = 2.0 * k * nb / (k + nb)
                # boundary faces (rolled-around neighbour) use the cell value
                sl = [slice(None)] * 3
                sl[ax_np] = -1 if st == 1 else 0
                h[tuple(sl)] = k[tuple(sl)]
                face_kappa[(ax, st)] = h
    return _grid_laplacian(shape, 3, face_k=face_kappa)


def diffusion_faces(nx):
    """7-point FV diffusion on an nx^3 grid parameterised by one conductance
    per face: 3 (nx-1) nx^2 interior faces, then 6 nx^2 Dirichlet boundary
    faces.  Returns (row_ptr, col_ind, m, n_faces, entry, face, sign): the
    matrix values for face conductances k are

        val = zeros(nnz); val[entry] += sign * k[face]

    i.e. a_pq = -k_f for the face f between cells p and q, and a_pp = sum of
    k_f over the six faces of cell p.  A(k) is SPD for k > 0.
    """
    n = nx
    m = n ** 3
    idx = np.arange(m, dtype=np.int64).reshape(n, n, n)  # [z, y, x]
    nin = (n - 1) * n * n
    nb = n * n
    # interior face ids per axis, on the lower cell of the pair
    ids = {}
    off = 0
    for ax_np, ax in ((2, 0), (1, 1), (0, 2)):
        shape = [n, n, n]
        shape[ax_np] = n - 1
        ids[ax] = off + np.arange(nin, dtype=np.int64).reshape(shape)
        off += nin
    bfirst = off
    n_faces = off + 6 * nb
    # face of every cell in every direction (interior or boundary)
    F = {}
    bk = 0
    for ax_np, ax in ((2, 0), (1, 1), (0, 2)):
        for st in (-1, 1):
            f = np.empty((n, n, n), dtype=np.int64)
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            if st == 1:
                lo[ax_np] = slice(0, n - 1)
                f[tuple(lo)] = ids[ax]
                hi[ax_np] = n - 1
                f[tuple(hi)] = bfirst + bk * nb + np.arange(nb).reshape(n, n)
            else:
                lo[ax_np] = slice(1, n)
                f[tuple(lo)] = ids[ax]
                hi[ax_np] = 0
                f[tuple(hi)] = bfirst + bk * nb + np.arange(nb).reshape(n, n)
            bk += 1
            F[(ax, st)] = f
    rp, ci, _, _ = _grid_laplacian((n, n, n), 3)
    # entry order inside a row: (2,-1) (1,-1) (0,-1) diag (0,1) (1,1) (2,1)
    order = [(2, -1), (1, -1), (0, -1), None, (0, 1), (1, 1), (2, 1)]
    coords = {0: np.broadcast_to(np.arange(n)[None, None, :], (n, n, n)),
              1: np.broadcast_to(np.arange(n)[None, :, None], (n, n, n)),
              2: np.broadcast_to(np.arange(n)[:, None, None], (n, n, n))}
    present = []
    for o in order:
        if o is None:
            present.append(np.ones((n, n, n), bool))
        else:
            c = coords[o[0]]
            present.append((c + o[1] >= 0) & (c + o[1] < n))
    P = np.stack([p.reshape(-1) for p in present], axis=1)
    epos = np.full(P.shape, -1, dtype=np.int64)
    epos[P] = np.arange(int(P.sum()))
    entry, face, sign = [], [], []
    for s, o in enumerate(order):
        if o is None:
            continue
        ok = P[:, s]
        entry.append(epos[ok, s])
        face.append(F[o].reshape(-1)[ok])
        sign.append(-np.ones(int(ok.sum())))
    dpos = epos[:, 3]
    for o in order:
        if o is None:
            continue
        entry.append(dpos)
        face.append(F[o].reshape(-1))
        sign.append(np.ones(m))
    return (rp, ci, m, n_faces, np.concatenate(entry), np.concatenate(face),
            np.concatenate(sign))


def _morton3(p, bits=10):
    """Z-order key of points in [0,1)^3 on a 2^bits grid per axis."""
    q = np.minimum((p * (1 << bits)).astype(np.int64), (1 << bits) - 1)
    key = np.zeros(len(p), dtype=np.int64)
    for b in range(bits):
        for a in range(3):
            key |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return key


def knn_laplacian(n_points, k=12, seed=0, order="morton"):
    """Weighted graph Laplacian of the symmetrised k-nearest-neighbour graph
    of n_points uniform points in [0,1]^3, parameterised by one weight per
    edge.  Returns (row_ptr, col_ind, m, edges, entry, edge, sign, diag_pos):
    for edge weights w (w > 0) and a shift s > 0 the SPD matrix L(w) + s I has

        val = zeros(nnz); val[entry] += sign * w[edge]; val[diag_pos] += s

    i.e. a_pq = -w_e for the edge e = (p, q) and a_pp = s + sum of w_e over
    the edges at p.  Columns are sorted within rows.  With order="morton"
    the points are numbered along a Z-order curve, so graph neighbours get
    nearby row numbers (the numbering a partitioner would produce); with
    "random" they keep the sampling order.
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.0, 1.0, (n_points, 3))
    if order == "morton":
        pts = pts[np.argsort(_morton3(pts), kind="stable")]
    elif order != "random":
        raise ValueError(order)
    _, nb = cKDTree(pts).query(pts, k=k + 1)
    i = np.repeat(np.arange(n_points, dtype=np.int64), k)
    j = nb[:, 1:].reshape(-1).astype(np.int64)
    keep = i != j
    lo, hi = np.minimum(i[keep], j[keep]), np.maximum(i[keep], j[keep])
    edges = np.unique(lo * n_points + hi)
    ep, eq = edges // n_points, edges % n_points
    ne = len(edges)
    m = n_points
    rows = np.concatenate([ep, eq, np.arange(m)])
    cols = np.concatenate([eq, ep, np.arange(m)])
    key = rows * m + cols
    srt = np.argsort(key, kind="stable")
    rows_s, cols_s = rows[srt], cols[srt]
    row_ptr = np.zeros(m + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum(np.bincount(rows_s, minlength=m))
    pos = np.empty(len(srt), dtype=np.int64)
    pos[srt] = np.arange(len(srt))
    e_ids = np.arange(ne, dtype=np.int64)
    dpos = pos[2 * ne:]
    entry = np.concatenate([pos[:ne], pos[ne:2 * ne], dpos[ep], dpos[eq]])
    edge = np.concatenate([e_ids, e_ids, e_ids, e_ids])
    sign = np.concatenate([-np.ones(2 * ne), np.ones(2 * ne)])
    return (row_ptr.astype(np.int32), cols_s.astype(np.int32), m,
            np.stack([ep, eq], axis=1), entry, edge, sign, dpos)


def assemble_knn(nnz, entry, edge, sign, diag_pos, w, shift=1e-2):
    val = assemble_faces(nnz, entry, edge, sign, w)
    val[diag_pos] += shift
    return val


def assemble_faces(nnz, entry, face, sign, k):
    """Values of A(k) for a diffusion_faces pattern (numpy)."""
    val = np.zeros(nnz)
    np.add.at(val, entry, sign * k[face])
    return val


def convection_diffusion2d(nx, wind=(20.0, -10.0)):
    """-lap u + wind . grad u on an nx x nx grid, Dirichlet boundaries by
    truncation; each convection term is upwinded (backward difference for a
    positive wind component, forward for a negative one), so off-diagonals
    are negative and rows are weakly diagonally dominant."""
    n = nx
    hh = 1.0 / (n + 1)
    m = n * n
    wx, wy = wind
    i = np.tile(np.arange(n), n)
    j = np.repeat(np.arange(n), n)
    d = 1.0 / hh**2
    # (column offset, coefficient, present) in ascending column order
    terms = [(-n, -d - max(wy, 0.0) / hh, j > 0),
             (-1, -d - max(wx, 0.0) / hh, i > 0),
             (0, 4.0 * d + abs(wx) / hh + abs(wy) / hh, np.ones(m, bool)),
             (1, -d + min(wx, 0.0) / hh, i < n - 1),
             (n, -d + min(wy, 0.0) / hh, j < n - 1)]
    counts = sum(t[2].astype(np.int64) for t in terms)
    row_ptr = np.zeros(m + 1, np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    col_ind = np.empty(row_ptr[-1], np.int32)
    val = np.empty(row_ptr[-1])
    pos = row_ptr[:-1].copy()
    rows = np.arange(m)
    for off, coef, present in terms:
        r = rows[present]
        col_ind[pos[r]] = r + off
        val[pos[r]] = coef
        pos[r] += 1
    return row_ptr.astype(np.int32), col_ind, val, m


def rhs_from_solution(row_ptr, col_ind, val, x):
    """b = A x in float64 (row-sequential sums, like the reference SpMV)."""
    m = len(row_ptr) - 1
    rows = np.repeat(np.arange(m), np.diff(row_ptr))
    prod = val * x[col_ind]
    return np.bincount(rows, weights=prod, minlength=m)


def uniform(n, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, n)


def normal(n, seed):
    return np.random.default_rng(seed).standard_normal(n)
