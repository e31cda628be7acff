"""CPU restatement of the learned sparse-approximate-inverse preconditioner.

TEST INFRASTRUCTURE: only tests, the smoke check and the benchmark's CPU
baseline use this module, as the checker of the device implementation.
Nothing in the product path imports it.

The reference solves A delta = g with UMFPACK (simulation.cpp:386-396) and
has no preconditioner; this model has no counterpart there.  Its arithmetic
is pinned by central finite differences of its own loss (tests), not by the
reference.

Model.  A is the Newton matrix (simulation.cpp:236-384) in its CSR view
(row i at Ap[i]..Ap[i+1]; A_ij = AxRow[k], A_ji = Ax[k] for j = Ai[k]; the
pattern is symmetric).  B^-1 is the pair-block inverse (mutual strongest
couplings as 2x2 blocks, else 1/A_ii), with diagonal bd and partner bcol.
The preconditioner is

    M = B^-1 + C,   C_ij = -gamma_ij * bd_i * A_ij * bd_j / (1 + |x_ij| + |w_ij|)
                    C_ii =  gamma_i  * bd_i * sum_j x_ij w_ij / (1 + |x_ij w_ij|)

with x_ij = bd_i A_ij, w_ij = bd_j A_ji (the damping keeps C bounded where a
first-order term is meaningless, |x| or |w| >> 1); the core edges are the off-diagonal
entries whose pair lies within the cutoff radius (r^2 <= rc^2, the untapered
part of the Jacobian), minus the pair-block partner.  gamma comes from a
two-layer message-passing network (hidden width 32) on the core graph:

    h0 = relu(W0 u_i + b0)
    layer l = 1, 2:
        a_ij = relu(P_l h_j + Q_l e_ij + c_l)
        s_i  = mean_j a_ij
        h_i  = relu(U_l h_i + V_l s_i + d_l)
    gamma_ij = wa.h_i + wb.h_j + wab.(h_i*h_j) + we.e_ij + beta
    gamma_i  = wd.h_i + beta_d

Loss on K Rademacher probes W (rows w_k): R = A M W^T - W^T, q_i = mean_k
R_ik^2, L = mean_i log(1 + q_i / eps): a per-row robust residual of the
right-preconditioned operator.
"""
import numpy as np

H = 32
FN = 4
FE = 4
EPS_LOSS = 1e-4

# flat parameter layout (float32 on the device)
LAYOUT = [("W0", (H, FN)), ("b0", (H,))]
for _l in (1, 2):
    LAYOUT += [(f"P{_l}", (H, H)), (f"Q{_l}", (H, FE)), (f"c{_l}", (H,)), (f"U{_l}", (H, H)), (f"V{_l}", (H, H)),
               (f"d{_l}", (H,))]
LAYOUT += [("wa", (H,)), ("wb", (H,)), ("wab", (H,)), ("we", (FE,)), ("beta", (1,)), ("wd", (H,)), ("betad", (1,))]
NPARAM = int(sum(np.prod(s) for _, s in LAYOUT))


def unpack(theta):
    out, o = {}, 0
    for name, shape in LAYOUT:
        size = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + size], dtype=np.float64).reshape(shape)
        o += size
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in LAYOUT])


MASK64 = (1 << 64) - 1


def splitmix_mix(z):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def init_params(seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from the probe hash stream;
    biases zero; readout zero (gamma = 0: the plain pair-block method)."""
    theta = np.zeros(NPARAM)
    o = 0
    for name, shape in LAYOUT:
        size = int(np.prod(shape))
        if len(shape) == 2:
            idx = np.arange(o, o + size, dtype=np.uint64)
            with np.errstate(over="ignore"):
                z = splitmix_mix(np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (idx + np.uint64(1)))
            u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
            theta[o:o + size] = (2.0 * u - 1.0) / np.sqrt(shape[1])
        o += size
    return theta.astype(np.float32).astype(np.float64)


def probes(seed, k, n):
    """Rademacher probe matrix (k x n): sign bit of the SplitMix64 hash of
    seed + golden * (p * n + i + 1)."""
    idx = (np.arange(k, dtype=np.uint64)[:, None] * np.uint64(n) + np.arange(n, dtype=np.uint64)[None, :])
    with np.errstate(over="ignore"):
        z = splitmix_mix(np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (idx + np.uint64(1)))
    return np.where((z >> np.uint64(63)) != 0, -1.0, 1.0)


def wrap(d):
    """utility.h:29-40 loop semantics for |d| < a few periods."""
    d = np.array(d, dtype=np.float64)
    for _ in range(8):
        d = np.where(d < -0.5, d + 1.0, d)
        d = np.where(d >= 0.5, d - 1.0, d)
    return d


class System:
    """One Newton matrix with its pair-block inverse and core graph."""

    def __init__(self, Ap, Ai, Ax, AxRow, diag, bd, bo, bcol, x, y, rc):
        n = len(Ap) - 1
        self.n = n
        self.Ap, self.Ai, self.Ax, self.AxRow, self.diag = Ap, Ai, Ax, AxRow, diag
        self.bd, self.bo, self.bcol = bd, bo, bcol
        rows = np.repeat(np.arange(n), np.diff(Ap))
        cols = Ai
        lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
        dx = wrap(x[hi] - x[lo])
        dy = wrap(y[hi] - y[lo])
        r2 = dx * dx + dy * dy
        core = (rows != cols) & ~(r2 > rc * rc) & (bcol[rows] != cols)
        self.k = np.nonzero(core)[0]                  # positions in A's CSR
        self.ri = rows[self.k]
        self.ci = cols[self.k]
        self.Cp = np.concatenate([[0], np.cumsum(np.bincount(self.ri, minlength=n))]).astype(np.int64)
        # mirror: edge index of (ci, ri)
        key = self.ri.astype(np.int64) * n + self.ci
        mkey = self.ci.astype(np.int64) * n + self.ri
        order = np.argsort(key)
        self.mirror = order[np.searchsorted(key[order], mkey)]
        assert np.array_equal(key[self.mirror], mkey)
        self.deg = np.diff(self.Cp)
        aij = AxRow[self.k]
        aji = Ax[self.k]
        xe = bd[self.ri] * aij
        ye = aij * bd[self.ci]
        we = bd[self.ci] * aji
        self.scale_e = -bd[self.ri] * aij * bd[self.ci] / (1.0 + np.abs(xe) + np.abs(we))
        t = np.zeros(n)
        np.add.at(t, self.ri, xe * we / (1.0 + np.abs(xe * we)))
        self.scale_d = bd * t
        sr = np.zeros(n)
        np.add.at(sr, self.ri, np.abs(xe))
        sc = np.zeros(n)
        np.add.at(sc, self.ri, np.abs(aji * bd[self.ri]))
        add = np.abs(AxRow[diag])
        rho = np.abs(add * bd)
        self.U = np.stack([sr / (1 + sr), sc / (1 + sc), (bcol >= 0).astype(float), rho / (1 + rho)], axis=1)
        xw = np.abs(xe * we)
        self.E = np.stack([xe / (1 + np.abs(xe)), ye / (1 + np.abs(ye)), we / (1 + np.abs(we)), xw / (1 + xw)],
                          axis=1)

    def csr_A(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.AxRow, self.Ai, self.Ap), shape=(self.n, self.n))

    def apply_M(self, Cx_e, Cx_d, V):
        """Z = M V for V (k x n)."""
        Z = self.bd[None, :] * V
        p = self.bcol >= 0
        Z[:, p] += self.bo[p][None, :] * V[:, self.bcol[p]]
        Z += Cx_d[None, :] * V
        contrib = Cx_e[None, :] * V[:, self.ci]
        for kk in range(V.shape[0]):
            Z[kk] += np.bincount(self.ri, weights=contrib[kk], minlength=self.n)
        return Z


def forward(theta, s: System):
    P = unpack(theta)
    c = {}
    a0 = s.U @ P["W0"].T + P["b0"]
    h = np.maximum(a0, 0)
    c["a0"], c["h0"] = a0, h
    deg = np.maximum(s.deg, 1)[:, None]
    for l in (1, 2):
        z = h @ P[f"P{l}"].T
        pe = z[s.ci] + s.E @ P[f"Q{l}"].T + P[f"c{l}"]
        ae = np.maximum(pe, 0)
        S = np.zeros((s.n, H))
        np.add.at(S, s.ri, ae)
        S /= deg
        ph = h @ P[f"U{l}"].T + S @ P[f"V{l}"].T + P[f"d{l}"]
        c[f"pe{l}"], c[f"S{l}"], c[f"ph{l}"], c[f"hin{l}"] = pe, S, ph, h
        h = np.maximum(ph, 0)
    c["h2"] = h
    hi, hj = h[s.ri], h[s.ci]
    ge = hi @ P["wa"] + hj @ P["wb"] + (hi * hj) @ P["wab"] + s.E @ P["we"] + P["beta"][0]
    gd = h @ P["wd"] + P["betad"][0]
    return ge, gd, c


def backward(theta, s: System, c, dge, dgd):
    P = unpack(theta)
    G = {name: np.zeros(shape) for name, shape in LAYOUT}
    h = c["h2"]
    hi, hj = h[s.ri], h[s.ci]
    G["wa"] = hi.T @ dge
    G["wb"] = hj.T @ dge
    G["wab"] = (hi * hj).T @ dge
    G["we"] = s.E.T @ dge
    G["beta"] = np.array([dge.sum()])
    G["wd"] = h.T @ dgd
    G["betad"] = np.array([dgd.sum()])
    dh = dgd[:, None] * P["wd"][None, :]
    ce = dge[:, None] * (P["wa"][None, :] + P["wab"][None, :] * hj)
    np.add.at(dh, s.ri, ce)
    cj = dge[:, None] * (P["wb"][None, :] + P["wab"][None, :] * hi)
    np.add.at(dh, s.ci, cj)
    deg = np.maximum(s.deg, 1)[:, None]
    for l in (2, 1):
        hin = c[f"hin{l}"]
        dph = dh * (c[f"ph{l}"] > 0)
        G[f"U{l}"] = dph.T @ hin
        G[f"V{l}"] = dph.T @ c[f"S{l}"]
        G[f"d{l}"] = dph.sum(0)
        dhin = dph @ P[f"U{l}"]
        dS = (dph @ P[f"V{l}"]) / deg
        dpe = dS[s.ri] * (c[f"pe{l}"] > 0)
        G[f"Q{l}"] = dpe.T @ s.E
        G[f"c{l}"] = dpe.sum(0)
        dz = np.zeros((s.n, H))
        np.add.at(dz, s.ci, dpe)
        G[f"P{l}"] = dz.T @ hin
        dhin += dz @ P[f"P{l}"]
        dh = dhin
    da0 = dh * (c["a0"] > 0)
    G["W0"] = da0.T @ s.U
    G["b0"] = da0.sum(0)
    return pack(G)


def loss_and_grad(theta, s: System, k, seed, eps=EPS_LOSS, with_grad=True):
    """Returns (L, dL/dtheta, gamma_e, gamma_d)."""
    ge, gd, c = forward(theta, s)
    Ce = ge * s.scale_e
    Cd = gd * s.scale_d
    W = probes(seed, k, s.n)
    A = s.csr_A()
    Y = s.apply_M(Ce, Cd, W)
    R = (A @ Y.T).T - W
    q = (R * R).mean(0)
    L = np.log1p(q / eps).mean()
    if not with_grad:
        return L, None, ge, gd
    # dL/dR_ki = (1/n) * 1/(eps+q_i) * 2 R_ki / k
    dR = (2.0 / (k * s.n)) * R / (eps + q)[None, :]
    AT = A.T.tocsr()
    dY = (AT @ dR.T).T                     # dL/dY = A^T dR
    dCe = (dY[:, s.ri] * W[:, s.ci]).sum(0)
    dCd = (dY * W).sum(0)
    g = backward(theta, s, c, dCe * s.scale_e, dCd * s.scale_d)
    return L, g, ge, gd


class Adam:
    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, theta, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return theta - self.lr * mh / (np.sqrt(vh) + self.eps)


def pair_block(Ap, Ai, Ax, AxRow, diag, threshold=1e-3):
    """Pair-block inverse as the device computes it (partner = argmax of
    |A_ij A_ji| / |A_ii A_jj|, ties to the smaller j; mutual partners above
    the threshold form 2x2 blocks)."""
    n = len(Ap) - 1
    add = AxRow[diag]
    partner = -np.ones(n, dtype=np.int64)
    strength = -np.ones(n)
    for i in range(n):
        ks = np.arange(Ap[i], Ap[i + 1])
        js = Ai[ks]
        sel = js != i
        if not sel.any():
            continue
        ks, js = ks[sel], js[sel]
        st = np.abs(AxRow[ks] * Ax[ks]) / (np.abs(add[i]) * np.abs(add[js]))
        best = st.max()
        partner[i] = js[st == best].min()
        strength[i] = best
    bd = 1.0 / add
    bo = np.zeros(n)
    bcol = -np.ones(n, dtype=np.int64)
    for i in range(n):
        j = partner[i]
        if j >= 0 and partner[j] == i and strength[i] > threshold:
            ks = np.arange(Ap[i], Ap[i + 1])
            kk = ks[Ai[ks] == j][0]
            aij, aji = AxRow[kk], Ax[kk]
            det = add[i] * add[j] - aij * aji
            bd[i] = add[j] / det
            bo[i] = -aij / det
            bcol[i] = j
    return bd, bo, bcol


def csr_view(Ap, Ai, Ax):
    """Values of A row-major on the CSC index arrays of a symmetric pattern:
    AxRow[k] = A[j, Ai[k]] for k in column j (the value stored at the mirror
    position)."""
    n = len(Ap) - 1
    cols = np.repeat(np.arange(n), np.diff(Ap))
    key = cols.astype(np.int64) * n + Ai          # (col, row) of each position
    mkey = Ai.astype(np.int64) * n + cols         # (row, col): the mirror
    order = np.argsort(key)
    pos = order[np.searchsorted(key[order], mkey)]
    assert np.array_equal(key[pos], mkey), "pattern is not symmetric"
    return Ax[pos]


def system_from_jacobian(x, y, Ap, Ai, Ax, diag, rc):
    """System of the Newton matrix in CSC form (Ap, Ai, Ax, diag positions)."""
    Ap = np.asarray(Ap, dtype=np.int64)
    Ai = np.asarray(Ai, dtype=np.int64)
    diag = np.asarray(diag, dtype=np.int64)
    AxRow = csr_view(Ap, Ai, Ax)
    bd, bo, bcol = pair_block(Ap, Ai, Ax, AxRow, diag)
    return System(Ap, Ai, Ax, AxRow, diag, bd, bo, bcol, np.asarray(x), np.asarray(y), rc)
