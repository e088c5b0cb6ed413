"""Shared spectral-element building blocks of the oracle (TEST INFRASTRUCTURE).

Restates, in NumPy/FP64:

* ``gll``            -- GLL nodes/weights (``flat_shell_engine/gll.m:1-39``): the reference's
                        literal table for n = 2..10, equal to the fixture extracted from
                        ``gll.m`` (``tests/golden/gll_table.json``) bit for bit and pinned by the
                        closed form (``gll_closed_form``: roots of P'_{n-1}, weights from
                        P_{n-1}) within a few ulps, in ``tests/test_oracle.py``.
* ``vandermonde``    -- ``flat_shell_engine/Vandermonde.m:1-12`` (Q = inv(V), V(i,j)=ksi_j^(i-1)).
* ``shape1d``        -- the 1D factors inside ``shape2D_prim.m:27-49`` / ``shape3D_prim_v2p.m``.
* ``shape2d_prim``   -- ``flat_shell_engine/shape2D_prim.m:27-58`` (Kronecker D(x)I, I(x)D with
                        the 1e-12 zeroing of ``:52,:54``); pinned by D*1=0, D*x=1 and the
                        closed-form Lagrange derivative matrix.
* ``shape3d_prim``   -- ``solid3D/shape3D_prim_v2p.m:1-72``.
* ``hanning_signal`` -- ``flat_shell_engine/Hanning_signal.m:17-23``; pinned by properties the
                        burst has whatever the form of its formula (support, carrier frequency,
                        antisymmetry about the centre, onset, energy; ``tests/test_oracle.py``).
* ``nodes_maps``     -- the 12-basket colouring ``flat_shell/mesher/nodesMaps_Fiborek.m:1-27``
                        (identical algorithm in ``solid3D/mesher3D/parallel_LG_nodes_Modified_Zb2.m``);
                        pinned by its invariants (unique nodes per basket column, bijection onto
                        all element-local points, balanced fills) and a hand-worked 2-element case.
* ``ChunkedForces``  -- the models' own ``local_forces`` run over contiguous element ranges on a
                        thread pool (the all-core leg of the CPU baseline); equal to the
                        one-range result bit for bit (``tests/test_oracle.py``).

Indices are 0-based here (the reference is 1-based); every translation is explicit.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as _leg


# gll.m:7-36: the reference's literal nodes and weights for n = 2..10, first ceil(n/2) of
# each; the rest by symmetry, as gll.m writes them (ksi(n+1-i) = -ksi(i), wi(n+1-i) = wi(i))
_GLL_TABLE = {
    2: ([-1.0], [1.0]),
    3: ([-1.0, 0.0], [0.3333333333333333, 1.3333333333333333]),
    4: ([-1.0, -0.4472135954999579], [0.16666666666666666, 0.8333333333333333]),
    5: ([-1.0, -0.6546536707079771, 0.0], [0.1, 0.5444444444444445, 0.7111111111111111]),
    6: ([-1.0, -0.7650553239294647, -0.28523151648064504],
        [0.06666666666666667, 0.37847495629784766, 0.5548583770354865]),
    7: ([-1.0, -0.8302238962785669, -0.4688487934707142, 0.0],
        [0.047619047619047616, 0.2768260473615661, 0.43174538120986283, 0.4876190476190476]),
    8: ([-1.0, -0.8717401485096066, -0.5917001814331423, -0.2092992179024788],
        [0.03571428571428571, 0.2107042271435067, 0.3411226924835044, 0.4124587946587037]),
    9: ([-1.0, -0.8997579954114602, -0.6771862795107377, -0.36311746382617816, 0.0],
        [0.027777777777777776, 0.16549536156080175, 0.27453871250016254, 0.34642851097304594,
         0.37151927437641724]),
    10: ([-1.0, -0.9195339081664589, -0.738773865105505, -0.4779249498104445, -0.16527895766638692],
         [0.022222222222222223, 0.1333059908510802, 0.22488934206312652, 0.29204268367968383,
          0.32753976118389744]),
}


def gll(n: int):
    """GLL nodes and weights on [-1, 1] for ``n`` points (``gll.m:1-39``): the reference's
    literal table for n = 2..10 (the values every reference computation starts from, to the
    last bit), the closed form (``gll_closed_form``) beyond it."""
    if n in _GLL_TABLE:
        h, wh = _GLL_TABLE[n]
        x = np.zeros(n)
        w = np.zeros(n)
        for i, (xi, wi) in enumerate(zip(h, wh)):
            x[i], w[i] = xi, wi
            if i < n // 2:  # the middle node of odd n stays +0
                x[n - 1 - i], w[n - 1 - i] = -xi, wi
        return x, w
    return gll_closed_form(n)


def gll_closed_form(n: int):
    """GLL nodes and weights from their definition: interior nodes are the roots of
    P'_{n-1}; w_i = 2 / (n (n-1) P_{n-1}(x_i)^2); made exactly antisymmetric (nodes) /
    symmetric (weights) as the reference's literal table is.  Pins that table (within a
    few ulps) and covers n > 10."""
    if n < 2:
        raise ValueError("gll needs n >= 2")
    N = n - 1
    cN = np.zeros(N + 1)
    cN[N] = 1.0
    if n == 2:
        x = np.array([-1.0, 1.0])
    else:
        inner = np.sort(np.real(_leg.legroots(_leg.legder(cN))))
        # one Newton polish on P'_N for full double accuracy
        d1 = _leg.legder(cN)
        d2 = _leg.legder(cN, 2)
        inner = inner - _leg.legval(inner, d1) / _leg.legval(inner, d2)
        x = np.concatenate([[-1.0], inner, [1.0]])
    x = 0.5 * (x - x[::-1])  # exact antisymmetry
    if n % 2 == 1:
        x[n // 2] = 0.0
    w = 2.0 / (N * (N + 1) * _leg.legval(x, cN) ** 2)
    w = 0.5 * (w + w[::-1])
    x[0], x[-1] = -1.0, 1.0
    return x, w


def vandermonde(ksi):
    """``Vandermonde.m:1-12``: V(i,j) = ksi_j^(i-1), Q = inv(V)."""
    ksi = np.asarray(ksi, dtype=np.float64)
    n = ksi.size
    V = np.empty((n, n))
    for i in range(n):
        V[i, :] = ksi ** i
    return np.linalg.inv(V)


def shape1d(Q, x):
    """1D factors of ``shape2D_prim.m:27-49``: S(k,j)=L_j(x_k), Sprim(k,j)=L_j'(x_k)."""
    x = np.asarray(x, dtype=np.float64)
    n = Q.shape[0]
    u = np.zeros((x.size, n))
    up = np.zeros((x.size, n))
    u[:, 0] = 1.0
    for j in range(1, n):
        u[:, j] = x ** j
        up[:, j] = j * x ** (j - 1)
    return u @ Q.T, up @ Q.T


def shape2d_prim(nx, ny, Qx, Qy, x, y):
    """``shape2D_prim.m:27-58``.  Local point c = k2*nx + k1 (xi fastest)."""
    Sx, Sxp = shape1d(Qx, x)
    Sy, Syp = shape1d(Qy, y)
    Nx = np.kron(Sy, Sxp)  # Nx[(k2,k1),(j2,j1)] = Sxp[k1,j1] * Sy[k2,j2]
    Ny = np.kron(Syp, Sx)
    Nx[np.abs(Nx) < 1e-12] = 0.0
    Ny[np.abs(Ny) < 1e-12] = 0.0
    return Nx, Ny


def shape3d_prim(nx, ny, nz, Qx, Qy, Qz, x, y, z):
    """``solid3D/shape3D_prim_v2p.m:1-72``.  Local point c = (k3*ny + k2)*nx + k1."""
    Sx, Sxp = shape1d(Qx, x)
    Sy, Syp = shape1d(Qy, y)
    Sz, Szp = shape1d(Qz, z)
    Nx = np.kron(Sz, np.kron(Sy, Sxp))
    Ny = np.kron(Sz, np.kron(Syp, Sx))
    Nz = np.kron(Szp, np.kron(Sy, Sx))
    for M in (Nx, Ny, Nz):
        M[np.abs(M) < 1e-12] = 0.0
    return Nx, Ny, Nz


def hanning_signal(dt, nft, f_1, f_2, t_1):
    """``Hanning_signal.m:17-23``: Hann-windowed tone burst, t(nf) = (nf-1)*dt."""
    t_2 = t_1 + 1.0 / f_1
    t = np.arange(nft, dtype=np.float64) * dt
    st = np.zeros(nft)
    m = (t >= t_1) & (t <= t_2)
    tm = t[m]
    st[m] = 0.5 * (1.0 - np.cos(2 * np.pi * f_1 * (tm - t_2))) * np.sin(2 * np.pi * f_2 * (tm - t_1))
    return t, st


def nodes_maps(nodes, n_nodes, n_baskets=12):
    """12-basket colouring, ``nodesMaps_Fiborek.m:1-27``.

    ``nodes`` is [fen x n_loc] (0-based node ids).  For each node k in ascending
    order its occurrences (element-major, then local index: ``:10-17``) go to the
    baskets of smallest current fill, ties to the lowest basket (MATLAB's stable
    ``sort``, ``:20``).  Returns 0-based ``I_G`` (node id) and ``I_L`` (element-local
    point e*n_loc + i), both [rows x 12], padded with -1 where a column is short.
    """
    nodes = np.asarray(nodes, dtype=np.int64)
    fen, nloc = nodes.shape
    L = fen * nloc
    flat = nodes.reshape(-1)
    # occurrence lists in (element, local) order: stable argsort by node id
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=n_nodes)
    if counts.max(initial=0) > n_baskets:
        raise ValueError("a node has more occurrences than baskets")
    start = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=start[1:])
    rows = -(-L // n_baskets)
    I_G = -np.ones((rows, n_baskets), dtype=np.int64)
    I_L = -np.ones((rows, n_baskets), dtype=np.int64)
    H = np.zeros(n_baskets, dtype=np.int64)
    for k in range(n_nodes):
        c = counts[k]
        if c == 0:
            continue
        S = np.argsort(H, kind="stable")
        occ = order[start[k]:start[k] + c]
        for i in range(c):
            b = S[i]
            r = H[b]
            H[b] += 1
            I_G[r, b] = k
            I_L[r, b] = occ[i]
    return I_G, I_L


def basket_assemble(I_G, I_L, f_loc, n_nodes, nf):
    """Assembly ``F(I_G_dofs(:,b)) += F2(I_L_dofs(:,b))`` for b = 1..12 in order
    (``main_flat_shell_force_excit_c.m:660-668``; DOF expansion
    ``global_local_12baskets_5dofs_v2.m:32-44``).  ``f_loc`` is [L x nf]."""
    F = np.zeros((n_nodes, nf))
    for b in range(I_G.shape[1]):
        m = I_G[:, b] >= 0
        F[I_G[m, b]] += f_loc[I_L[m, b]]
    return F.reshape(-1)



def _chunk_worker(view, lo, hi, u_buf, f_buf, nf, start, done, stop):
    U = np.frombuffer(u_buf, dtype=np.float64)
    F = np.frombuffer(f_buf, dtype=np.float64).reshape(-1, nf)
    while True:
        start.wait()
        if stop.value:
            return
        F[lo:hi] = view.local_forces(U)
        done.wait()


class ChunkedForces:
    """``model.local_forces`` over ``n_chunks`` contiguous element ranges, one forked worker
    process per range (the all-core leg of the CPU baseline).

    Each range runs the model's own expressions on a view of the model whose per-point
    arrays are slices and whose block-diagonal derivative matrices hold only that range's
    (identical) element blocks, so every point's arithmetic is the same as in the one-range
    product: the result is bitwise that of ``model.local_forces``.  U and the forces live in
    shared memory; two barriers per call start the workers and collect them."""

    def __init__(self, model, n_chunks):
        import copy
        import multiprocessing as mp
        import scipy.sparse as sp
        ctx = mp.get_context("fork")
        L = model.fen * model.nloc
        self.nf = model.nf
        bounds = np.linspace(0, model.fen, max(1, int(n_chunks)) + 1).astype(np.int64)
        views, blocks = [], {}
        for lo, hi in zip(bounds[:-1], bounds[1:]):
            if hi <= lo:
                continue
            v = copy.copy(model)
            sl = slice(int(lo) * model.nloc, int(hi) * model.nloc)
            for name, val in vars(model).items():
                if name in ("mass", "dv", "nodes", "chunked"):
                    continue
                if isinstance(val, np.ndarray) and val.ndim == 1 and val.shape[0] == L:
                    setattr(v, name, val[sl])
                elif sp.issparse(val) and val.shape == (L, L):
                    key = (name, int(hi - lo))
                    if key not in blocks:
                        blk = sp.csr_matrix(val[:model.nloc, :model.nloc])
                        blocks[key] = sp.block_diag([blk] * int(hi - lo), format="csr")
                    setattr(v, name, blocks[key])
            v.chunked = None
            views.append((v, sl.start, sl.stop))
        self.u_buf = ctx.RawArray("d", model.dof)
        self.f_buf = ctx.RawArray("d", L * model.nf)
        self.U = np.frombuffer(self.u_buf, dtype=np.float64)
        self.F = np.frombuffer(self.f_buf, dtype=np.float64).reshape(L, model.nf)
        self.start = ctx.Barrier(len(views) + 1)
        self.done = ctx.Barrier(len(views) + 1)
        self.stop = ctx.RawValue("i", 0)
        self.procs = [ctx.Process(target=_chunk_worker, daemon=True,
                                  args=(v, lo, hi, self.u_buf, self.f_buf, model.nf, self.start, self.done,
                                        self.stop)) for v, lo, hi in views]
        for p in self.procs:
            p.start()

    def __call__(self, U):
        self.U[:] = U
        self.start.wait()
        self.done.wait()
        return self.F.copy()

    def close(self):
        self.stop.value = 1
        self.start.wait()
        for p in self.procs:
            p.join(10)
