"""Synthetic inputs of the benchmark and test configurations (host side, numpy + scipy).

The meshes follow the structured fixtures of packages/physics/src/common/mesh.ts: rectTriMesh (:69-121, alternating
diagonals) and boxMesh (:123-190, 6 Kuhn tets per hex sharing the 0-6 diagonal, orientation fixed by the volume sign).
Matrices are P1 stiffness matrices sum_e k_e |T_e| grad(phi_a) . grad(phi_b) on the full symmetric CSR pattern of
buildElementPattern (sparse.mjs:86-110: every row sorted, diagonal present), restricted to the free DOFs as restrict()
does (scalar.ts:210-233). Random draws use the reference's LCG s <- (1664525 s + 1013904223) mod 2^32, u = s / 2^32
(la-real-solve.test.mjs:33-34, eigen.ts:174).
"""
import numpy as np
import scipy.sparse as sp

LCG_A = 1664525
LCG_C = 1013904223
M32 = 1 << 32


def lcg_u32(seed, count):
    """The first `count` LCG states after `seed` (uint64 array holding 32-bit values), vectorised by block jumps."""
    count = int(count)
    out = np.empty(count, dtype=np.uint64)
    if count == 0:
        return out
    blk = min(count, 4096)
    s = int(seed) % M32
    for i in range(blk):
        s = (LCG_A * s + LCG_C) % M32
        out[i] = s
    ja, jc = 1, 0  # affine map of `blk` steps
    for _ in range(blk):
        ja, jc = (LCG_A * ja) % M32, (LCG_A * jc + LCG_C) % M32
    ja64, jc64, m64 = np.uint64(ja), np.uint64(jc), np.uint64(M32)
    i = blk
    while i < count:
        k = min(blk, count - i)
        out[i:i + k] = ((ja64 * out[i - blk:i - blk + k]) % m64 + jc64) % m64
        i += k
    return out


def lcg_uniform(seed, count):
    """u = s / 2^32 in [0, 1)."""
    return lcg_u32(seed, count).astype(np.float64) / float(M32)


def lcg_normal(seed, count):
    """Box-Muller on consecutive LCG pairs (u1 shifted off zero)."""
    m = (count + 1) // 2
    u = lcg_uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return z[:count]


# ---- meshes -----------------------------------------------------------------------------------------------------

def rect_tri_mesh(nx, ny):
    """Unit square, nx x ny cells, 2 P1 triangles per cell with alternating diagonals (mesh.ts:69-91)."""
    NX, NY = nx, ny
    j, i = np.meshgrid(np.arange(NY + 1), np.arange(NX + 1), indexing="ij")
    nodes = np.stack([i.ravel() / NX, j.ravel() / NY], axis=1).astype(np.float64)
    nid = lambda a, b: b * (NX + 1) + a  # noqa: E731
    cj, ci = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ci, cj = ci.ravel(), cj.ravel()
    flip = (ci + cj) % 2 == 1
    c00, c10, c11, c01 = nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)
    t1 = np.where(flip[:, None], np.stack([c00, c10, c11], 1), np.stack([c00, c10, c01], 1))
    t2 = np.where(flip[:, None], np.stack([c00, c11, c01], 1), np.stack([c10, c11, c01], 1))
    tris = np.empty((2 * len(ci), 3), dtype=np.int64)
    tris[0::2] = t1
    tris[1::2] = t2
    return nodes, tris


_HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
_KUHN = [[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]]


def _kuhn_corners():
    """Corner offsets of the 6 Kuhn tets, reordered to positive volume (mesh.ts:138-151)."""
    out = []
    for t in _KUHN:
        cs = _HEX[t].astype(np.float64)
        u, v, w = cs[1] - cs[0], cs[2] - cs[0], cs[3] - cs[0]
        if np.dot(u, np.cross(v, w)) < 0:
            cs = cs[[0, 2, 1, 3]]
        out.append(cs.astype(np.int64))
    return np.stack(out)  # (6, 4, 3)


def box_tet_mesh(nx, ny, nz):
    """Unit cube, nx x ny x nz hexes, 6 Kuhn tets each (mesh.ts:123-150). Tets ordered hex by hex (k, j, i)."""
    NX, NY, NZ = nx, ny, nz
    kk, jj, ii = np.meshgrid(np.arange(NZ + 1), np.arange(NY + 1), np.arange(NX + 1), indexing="ij")
    nodes = np.stack([ii.ravel() / NX, jj.ravel() / NY, kk.ravel() / NZ], 1).astype(np.float64)
    hk, hj, hi = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([hi.ravel(), hj.ravel(), hk.ravel()], 1)  # (nh, 3)
    kc = _kuhn_corners()  # (6, 4, 3)
    pts = base[:, None, None, :] + kc[None]  # (nh, 6, 4, 3)
    ids = (pts[..., 2] * (NY + 1) + pts[..., 1]) * (NX + 1) + pts[..., 0]
    return nodes, ids.reshape(-1, 4)


# ---- P1 assembly ---------------------------------------------------------------------------------------------

def p1_element_stiffness(coords):
    """Geometric P1 stiffness |T| G G^T for simplices given as (ne, d+1, d) coordinates."""
    d = coords.shape[2]
    J = np.transpose(coords[:, 1:, :] - coords[:, :1, :], (0, 2, 1))  # columns are edge vectors
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)  # rows: gradients of barycentrics 1..d
    G = np.concatenate([-Jinv.sum(axis=1, keepdims=True), Jinv], axis=1)  # (ne, d+1, d)
    vol = np.abs(det) / (2.0 if d == 2 else 6.0)
    return vol[:, None, None] * np.einsum("eai,ebi->eab", G, G)


def assemble(n, conn, Ke):
    """Full symmetric CSR (sorted rows, diagonal present) of element matrices Ke (ne, k, k) on conn (ne, k): the
    pattern of buildElementPattern (sparse.mjs:86-110) with the element entries summed into their slots."""
    k = conn.shape[1]
    rows = np.repeat(conn, k, axis=1).ravel().astype(np.int64)
    cols = np.tile(conn, (1, k)).ravel().astype(np.int64)
    keys = np.concatenate([rows * n + cols, np.arange(n, dtype=np.int64) * (n + 1)])
    ukeys = np.unique(keys)
    slot = np.searchsorted(ukeys, rows * n + cols)
    data = np.bincount(slot, weights=Ke.ravel(), minlength=len(ukeys))
    r, c = ukeys // n, ukeys % n
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=indptr[1:])
    return sp.csr_matrix((data, c.astype(np.int32), indptr), shape=(n, n))


def restrict(A, free_mask):
    """Rows / columns with free_mask, order kept (scalar.ts:210-233)."""
    idx = np.nonzero(free_mask)[0]
    R = A[idx][:, idx].tocsr()
    R.sort_indices()
    return R, idx


def csr_dict(A):
    A = A.tocsr()
    A.sort_indices()
    return {"n": A.shape[0], "rowptr": A.indptr.astype(np.int32), "colidx": A.indices.astype(np.int32),
            "nnz": int(A.nnz)}


def poisson2d(nx=64, seed=1):
    """Config: 2D Poisson, P1 triangles on the unit square, nx x nx cells, Dirichlet on every edge, k = 1,
    RHS U[-1, 1] from the LCG. Returns (pat, values, b)."""
    nodes, tris = rect_tri_mesh(nx, nx)
    Ke = p1_element_stiffness(nodes[tris])
    A = assemble(len(nodes), tris, Ke)
    x, y = nodes[:, 0], nodes[:, 1]
    eps = 1e-12
    free = (x > eps) & (x < 1 - eps) & (y > eps) & (y < 1 - eps)
    R, _ = restrict(A, free)
    b = 2.0 * lcg_uniform(seed, R.shape[0]) - 1.0
    return csr_dict(R), R.data.astype(np.float64), b


def heat3d(ncell=60, seed=1, rhs_seed=2):
    """Config: 3D steady heat, P1 tets on the unit cube, ncell^3 hexes x 6 Kuhn tets, per-element conductivity
    10^(2u) (log-uniform in [1, 100]) from the LCG, Dirichlet on z = 0. Returns (pat, values, b)."""
    nodes, tets = box_tet_mesh(ncell, ncell, ncell)
    kc = _kuhn_corners().astype(np.float64) / ncell
    Kref = p1_element_stiffness(kc)  # the 6 Kuhn shapes: every hex has the same geometry
    kel = 10.0 ** (2.0 * lcg_uniform(seed, len(tets)))
    Ke = Kref[np.arange(len(tets)) % 6] * kel[:, None, None]
    A = assemble(len(nodes), tets, Ke)
    free = nodes[:, 2] > 1e-12
    R, _ = restrict(A, free)
    b = 2.0 * lcg_uniform(rhs_seed, R.shape[0]) - 1.0
    return csr_dict(R), R.data.astype(np.float64), b


_TET_EDGES = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]])


def _maxwell_elements(ncell, seed, eps_choices):
    """Whitney edge-element matrices on the unit cube (ncell^3 hexes x 6 Kuhn tets): per-tet curl-curl Ke, unit mass
    Me, eps_r per tet, the signs of the global edge orientation (lower to higher node index) per local pair, the tets'
    global edge ids, and the mask of edges on the boundary (PEC)."""
    nodes, tets = box_tet_mesh(ncell, ncell, ncell)
    X = nodes[tets]  # (ne, 4, 3)
    J = np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))
    vol = np.abs(np.linalg.det(J)) / 6.0
    Jinv = np.linalg.inv(J)
    G = np.concatenate([-Jinv.sum(axis=1, keepdims=True), Jinv], axis=1)  # (ne, 4, 3) grad(lambda)
    ne = len(tets)
    nch = len(eps_choices)
    eps = np.array(eps_choices)[np.minimum((nch * lcg_uniform(seed, ne)).astype(np.int64), nch - 1)]
    I, Jn = _TET_EDGES[:, 0], _TET_EDGES[:, 1]
    curl = 2.0 * np.cross(G[:, I, :], G[:, Jn, :])  # (ne, 6, 3)
    Ke = vol[:, None, None] * np.einsum("eai,ebi->eab", curl, curl)
    GG = np.einsum("eai,ebi->eab", G, G)  # grad l_p . grad l_q
    d = lambda p, q: (1.0 + (p[:, None] == q[None, :]))  # noqa: E731
    i_, j_ = I, Jn
    Me = (vol / 20.0)[:, None, None] * (
        d(i_, i_) * GG[:, j_][:, :, j_] - d(i_, j_) * GG[:, j_][:, :, i_]
        - d(j_, i_) * GG[:, i_][:, :, j_] + d(j_, j_) * GG[:, i_][:, :, i_])
    a, b = tets[:, I], tets[:, Jn]  # local edge endpoints (global node ids)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    sgn = np.where(a < b, 1.0, -1.0)
    nn = len(nodes)
    keys = lo * nn + hi
    ukeys, eid = np.unique(keys, return_inverse=True)
    eid = eid.reshape(ne, 6)
    ss = sgn[:, :, None] * sgn[:, None, :]
    e0, e1 = ukeys // nn, ukeys % nn
    on = np.zeros(len(ukeys), dtype=bool)
    for ax in range(3):
        for v in (0.0, 1.0):
            on |= (np.abs(nodes[e0, ax] - v) < 1e-12) & (np.abs(nodes[e1, ax] - v) < 1e-12)
    return Ke, Me, eps, ss, eid, len(ukeys), on


def maxwell_edge(ncell=20, sigma=50.0, seed=3, rhs_seed=4, eps_choices=(1.0, 2.1, 11.7)):
    """Config: shift-inverted 3D Maxwell, lowest-order Nedelec (Whitney) edge elements on the unit cube, ncell^3 hexes
    x 6 Kuhn tets, PEC walls eliminated, per-element eps_r drawn uniformly from {1.0, 2.1, 11.7}, mu_r = 1; the matrix
    is K - sigma M (symmetric indefinite). Element matrices: curl w_ij = 2 grad(l_i) x grad(l_j),
    int w_a . w_b from int l_p l_q = |T| (1 + delta_pq) / 20 (as fem/nedelec1.mjs assembles them). Edges are oriented
    from the lower to the higher node index. Returns (pat, values, b)."""
    Ke, Me, eps, ss, eid, nedge, on = _maxwell_elements(ncell, seed, eps_choices)
    A = assemble(nedge, eid, (Ke - sigma * eps[:, None, None] * Me) * ss)
    R, _ = restrict(A, ~on)
    b = lcg_normal(rhs_seed, R.shape[0])
    return csr_dict(R), R.data.astype(np.float64), b


def maxwell_km(ncell=8, seed=3, eps_choices=(1.0,)):
    """The curl-curl stiffness K and the eps_r-weighted mass M of maxwell_edge on one pattern (PEC eliminated), for
    the eigenproblem K x = k^2 M x. Returns (pat, K values, M values)."""
    Ke, Me, eps, ss, eid, nedge, on = _maxwell_elements(ncell, seed, eps_choices)
    K, _ = restrict(assemble(nedge, eid, Ke * ss), ~on)
    M, _ = restrict(assemble(nedge, eid, eps[:, None, None] * Me * ss), ~on)
    assert np.array_equal(K.indptr, M.indptr) and np.array_equal(K.indices, M.indices)
    return csr_dict(K), K.data.astype(np.float64), M.data.astype(np.float64)


def heat2d(nx, kel, Dirichlet=True):
    """2D P1 heat on the unit square, nx x nx cells split into triangles (rect_tri_mesh), per-triangle conductivity
    kel (len 2 nx^2), Dirichlet on every edge. Returns the restricted CSR matrix."""
    nodes, tris = rect_tri_mesh(nx, nx)
    Ke = p1_element_stiffness(nodes[tris]) * np.asarray(kel, dtype=np.float64)[:, None, None]
    A = assemble(len(nodes), tris, Ke)
    x, y = nodes[:, 0], nodes[:, 1]
    eps = 1e-12
    free = (x > eps) & (x < 1 - eps) & (y > eps) & (y < 1 - eps)
    R, _ = restrict(A, free)
    return R


def batch_heat2d(nsys=512, lo=20, hi=120, nrhs=8, seed=5, rhs_seed=6):
    """Config: nsys independent 2D P1 heat systems on structured unit-square grids of side s ~ U{lo..hi} cells (LCG),
    per-triangle conductivity log-uniform in [1, 100] (10^(2u), one LCG stream over all systems in order), Dirichlet
    on every edge, nrhs normal right-hand sides each. The systems are assembled into one block-diagonal pattern
    (system i at rows offsets[i] .. offsets[i + 1]); right-hand side k of the union stacks the systems' k-th
    right-hand sides. Returns (pat, values, B (nrhs x n), sides, offsets)."""
    sides = lo + np.minimum((lcg_uniform(seed, nsys) * (hi - lo + 1)).astype(np.int64), hi - lo)
    ntri = 2 * sides ** 2
    kel = 10.0 ** (2.0 * lcg_uniform(seed + 1, int(ntri.sum())))
    kofs = np.concatenate([[0], np.cumsum(ntri)])
    blocks = [heat2d(int(s), kel[kofs[i]:kofs[i + 1]]) for i, s in enumerate(sides)]
    sizes = np.array([b.shape[0] for b in blocks], dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    A = sp.block_diag(blocks, format="csr")
    A.sort_indices()
    n = int(offsets[-1])
    B = lcg_normal(rhs_seed, nrhs * n).reshape(nrhs, n)
    return csr_dict(A), A.data.astype(np.float64), B, sides, offsets
