"""Test helpers: distorted UV maps, a grid air mesh for the scaffold term,
and the oracle-side composition of the scaffold objective
(Optimizer.cpp:764-843, Scaffold.cpp:210-313)."""
import numpy as np

from oracle import oracle as O


def distort(UV, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    k = rng.normal(size=4) * amp
    out = UV + np.stack([k[0] * np.sin(3 * UV[:, 1]) + k[1] * UV[:, 0] ** 2,
                         k[2] * np.cos(2 * UV[:, 0]) + k[3] * UV[:, 1] ** 2], 1)
    return out


def grid_air_mesh(m, rings=2):
    """Air mesh around the m x m unit-square grid of height_field_disc.

    The air region is the band of `rings` grid cells around [0,1]^2, split
    into CCW triangles. Returns (airUV, Fair_local, bnd) where the first
    len(bnd) air vertices are the mesh boundary vertices bnd (mesh ids).
    """
    s = m - 1
    h = 1.0 / s
    M = m + 2 * rings
    xs = (np.arange(M) - rings) * h
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    P = np.stack([X.ravel(), Y.ravel()], 1)
    big_id = lambda i, j: j * M + i  # noqa: E731
    # mesh vertex id of big-grid node inside the unit square
    def mesh_id(i, j):
        ii, jj = i - rings, j - rings
        if 0 <= ii < m and 0 <= jj < m:
            return jj * m + ii
        return -1
    tris = []
    for j in range(M - 1):
        for i in range(M - 1):
            inside = (rings <= i < rings + s) and (rings <= j < rings + s)
            if inside:
                continue
            a, b, c, d = big_id(i, j), big_id(i + 1, j), big_id(i, j + 1), big_id(i + 1, j + 1)
            tris.append((a, b, d))
            tris.append((a, d, c))
    tris = np.array(tris)
    used = np.unique(tris)
    # order: boundary (mesh) vertices first, then the rest
    bnd_big = [v for v in used if mesh_id(v % M, v // M) >= 0]
    rest_big = [v for v in used if mesh_id(v % M, v // M) < 0]
    order = bnd_big + rest_big
    remap = {v: k for k, v in enumerate(order)}
    Fair = np.vectorize(remap.get)(tris).astype(np.int32)
    airUV = P[order]
    bnd = np.array([mesh_id(v % M, v // M) for v in bnd_big], np.int32)
    return airUV, Fair, bnd


def flip_air_quads(air, seed, frac=0.5):
    """The same air vertices triangulated differently: grid_air_mesh emits each
    quad as (a,b,d), (a,d,c); a random subset takes the other diagonal."""
    airUV, Fair, bnd = air
    rng = np.random.default_rng(seed)
    F = Fair.copy()
    for q in np.flatnonzero(rng.random(F.shape[0] // 2) < frac):
        a, b, d = Fair[2 * q]
        c = Fair[2 * q + 1, 2]
        F[2 * q] = (a, b, c)
        F[2 * q + 1] = (b, d, c)
    return airUV, F, bnd


class ScaffoldOracle:
    """Objective a0*E_SD(mesh) + (w_scaf/F_air)*sum SD_uniform(air) on the oracle."""

    def __init__(self, Vrest, F, airUV, Fair, bnd, area_thres, a0, w_scaf):
        self.F, self.Fair, self.bnd = F, Fair, bnd
        self.nV = Vrest.shape[0]
        self.nVair = airUV.shape[0]
        self.nB = len(bnd)
        self.a0, self.sAir = a0, w_scaf / Fair.shape[0]
        self.feat = O.RestFeatures(Vrest, F)
        Vr = np.concatenate([airUV, np.zeros((self.nVair, 1))], 1)
        self.afeat = O.RestFeatures(Vr, Fair, area_thres=area_thres)
        self.l2g = np.concatenate([bnd, self.nV + np.arange(self.nVair - self.nB)]).astype(np.int64)
        self.air_fixed = [k for k in range(self.nB) if bnd[k] == 0]

    def split(self, X):
        """Combined UV (nV + nVair - nB, 2) -> (mesh UV, air UV)."""
        return X[: self.nV], X[self.l2g]

    def energy(self, X):
        U, A = self.split(X)
        em = O.energy(self.F, U, self.feat)
        ea = O.energy(self.Fair, A, self.afeat, uniform=True)
        return self.a0 * em + self.sAir * ea, em, self.sAir * ea

    def gradient(self, X):
        U, A = self.split(X)
        g = np.zeros(2 * X.shape[0])
        g[: 2 * self.nV] = self.a0 * O.gradient(self.F, U, self.feat, fixed=(0,))
        ga = O.gradient(self.Fair, A, self.afeat, fixed=tuple(self.air_fixed), uniform=True)
        for l in range(self.nVair):
            G = self.l2g[l]
            g[2 * G: 2 * G + 2] += self.sAir * ga[2 * l: 2 * l + 2]
        return g

    def hessian_upper(self, X, project=True):
        """Combined upper CSR (ia, ja, a) as Optimizer::computeHessian + update_a build it (project=False keeps
        the element Hessians unprojected, for derivative checks)."""
        U, A = self.split(X)
        I1, J1, V1 = O.hessian_triplets(self.F, O.element_hessians(self.F, U, self.feat, project=project),
                                        fixed=(0,), scale=self.a0)
        Ha = O.element_hessians(self.Fair, A, self.afeat, uniform=True, project=project)
        I2, J2, V2 = O.hessian_triplets(self.Fair, Ha, fixed=tuple(self.air_fixed), scale=self.sAir)
        I2 = (self.l2g[I2 // 2] * 2 + I2 % 2).astype(np.int32)
        J2 = (self.l2g[J2 // 2] * 2 + J2 % 2).astype(np.int32)
        nVt = self.nV + self.nVair - self.nB
        Fall = np.concatenate([self.F, self.l2g[self.Fair]]).astype(np.int32)
        ptr, nbr = O.vertex_neighbors(nVt, Fall)
        ia, ja = O.set_pattern(nVt, ptr, nbr, fixed=(0,))
        a = O.update_a(ia, ja, np.concatenate([I1, I2]), np.concatenate([J1, J2]), np.concatenate([V1, V2]))
        return ia, ja, a

    def newton_step(self, X, lastE):
        """One scaffolded Optimizer::solve(1) iteration with the oracle's LDLT."""
        g = self.gradient(X)
        ia, ja, a = self.hessian_upper(X)
        d = O.LDLT(ia, ja, a).solve(-g)
        U, A = self.split(X)
        D = d.reshape(-1, 2)
        step = O.init_step_size(self.F, U, D[: self.nV].reshape(-1), 1.0)
        step = O.init_step_size(self.Fair, A, D[self.l2g].reshape(-1), step)
        step *= 0.99
        lastE = self.energy(X)[0]
        E = self.energy(X + step * D)[0]
        halv = 0
        while E > lastE:
            step /= 2.0
            halv += 1
            E = self.energy(X + step * D)[0]
        return X + step * D, E, step, halv, float(g @ g)


def block_system(V, F, UV, fixed=(0,)):
    """Newton matrix of the mesh-only objective as the full symmetric 2x2-block
    CSR the device driver assembles: fixed rows keep an identity diagonal
    block only, free rows drop fixed columns. Returns (rowPtr, colIdx, vals
    (nnzb x 2 x 2), fixedMask)."""
    import scipy.sparse as sp
    nV = V.shape[0]
    feat = O.RestFeatures(V, F)
    H = O.element_hessians(F, UV, feat)
    I, J, Vt = O.hessian_triplets(F, H)
    A = sp.coo_matrix((Vt, (I, J)), shape=(2 * nV, 2 * nV)).tocsr()
    fm = np.zeros(nV, np.uint8)
    fm[list(fixed)] = 1
    nbr = [set() for _ in range(nV)]
    for t in F:
        for a in t:
            for b in t:
                nbr[a].add(int(b))
    rp, ci, vals = [0], [], []
    A = A.tolil()
    for u in range(nV):
        cols = [u] if fm[u] else sorted(w for w in nbr[u] if w == u or not fm[w])
        for w in cols:
            ci.append(w)
            if fm[u]:
                vals.append(np.eye(2))
            else:
                vals.append(A[2 * u:2 * u + 2, 2 * w:2 * w + 2].toarray())
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(vals), fm


def block_csr_to_sparse(rp, ci, vals):
    import scipy.sparse as sp
    nV = len(rp) - 1
    row = np.repeat(np.arange(nV), np.diff(rp))
    r = (2 * row[:, None, None] + np.arange(2)[None, :, None]).repeat(2, 2)
    c = (2 * ci[:, None, None] + np.arange(2)[None, None, :]).repeat(2, 1)
    return sp.csr_matrix((np.asarray(vals).ravel(), (r.ravel(), c.ravel())), shape=(2 * nV, 2 * nV))


def ic0_dense(A, P):
    """Sequential incomplete Cholesky IC(0) of the dense symmetric matrix A (2n x 2n) on the 2x2-block pattern P
    (n x n boolean, symmetric, with its diagonal): the lower factor L, zero outside the lower half of the expanded
    pattern, with (L L^T)_ij = A_ij on it (the defining property of IC(0)). Row by row, left to right:
    L_ij = (A_ij - sum_{k<j} L_ik L_jk) / L_jj and L_ii = sqrt(A_ii - sum_{k<i} L_ik^2); L is zero off the
    pattern, so the sums run over the columns both rows hold. Raises ValueError on a non-positive pivot."""
    n = P.shape[0]
    S = np.kron(np.asarray(P, bool), np.ones((2, 2), bool))
    N = 2 * n
    L = np.zeros((N, N))
    for i in range(N):
        for j in np.nonzero(S[i, : i + 1])[0]:
            s = A[i, j] - L[i, :j] @ L[j, :j]
            if i == j:
                if not s > 0.0:
                    raise ValueError("IC(0) breakdown at row %d" % i)
                L[i, i] = np.sqrt(s)
            else:
                L[i, j] = s / L[j, j]
    return L


def upper_csr_to_dense(ia, ja, a):
    """Full symmetric dense matrix of an upper-triangle CSR (the reference's LinSysSolver ia/ja/a)."""
    N = len(ia) - 1
    U = np.zeros((N, N))
    for i in range(N):
        for k in range(ia[i], ia[i + 1]):
            U[i, ja[k]] = a[k]
    return U + U.T - np.diag(np.diag(U))


def block_pattern_of_upper(ia, ja):
    """2x2-block pattern (n x n boolean, symmetric, with the diagonal) of an upper-triangle DOF CSR."""
    N = len(ia) - 1
    P = np.eye(N // 2, dtype=bool)
    for i in range(N):
        for k in range(ia[i], ia[i + 1]):
            P[i // 2, ja[k] // 2] = P[ja[k] // 2, i // 2] = True
    return P
