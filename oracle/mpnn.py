"""FP64 CPU restatement of the learned stepper (test infrastructure).

Only tests, smoke() and the bench's CPU leg import this module; the library
never does.  The learned stepper has no counterpart in the reference
(mfree_iwf integrates with the RK4 of src/rk4.h:71-108 only), so this file
restates the project's own definitions, in float64 numpy:

  radius_graph   edges r < 2 h_ij, h_ij = (h_i + h_j)/2, self excluded;
                 w = W(r, h_ij) h_ij^2, (gx, gy) = grad_i W(r, h_ij) h_ij^3
                 with the reference's cubic spline (src/kernel.cpp:20-55)
  forward        3 message-passing layers, hidden 128, SiLU, residual
  backward       exact gradients of the mean-squared loss (the transposed
                 aggregation is a scatter over the edges, it does not assume
                 a symmetric graph)
  rk_stages      the base step, the learned step without the network: one
                 step of the classical RK4 (src/rk4.h:71-108) or of Kutta's
                 third-order method over tau of a derivative function, with
                 its stage rates
  features       the 16 node inputs of a learned step (start state, previous
                 velocity, the base step's stage differences)
  conserve       the correction's projection onto zero mass-weighted mean
                 of x, y, vx, vy (linear momentum and centre of mass)

Pins (tests/test_mpnn_oracle.py): finite differences of the loss against
backward(); torch.autograd in float64 against forward() and backward(); a
per-edge Python loop against the vectorised aggregation and the adjoint
identity of its transpose; radius_graph's weights against the
reference-pinned cubic spline of the CPU oracle; the graph's symmetry;
rk_stages against the degree-3 / degree-4 Taylor polynomial of exp(tau A)
for a linear system y' = A y (the stability polynomial, which pins every
coefficient of the tableau) and against the exact solution of uniformly
accelerated motion; features against their invariance under a change of
units and the stage differences against (tau A)^p y of a linear system;
targets and correction as inverse maps; conserve as the orthogonal
projection of the mass-weighted inner product (idempotent, zero sums,
closest point of the subspace).
"""
import numpy as np

NF = 16
K0 = 32
H = 128
C = 32
G = H + 3 * C
NOUT = 8
LAYERS = 3
SPLINE_NORM = 10.0 / (7.0 * np.pi)


def layout():
    shapes = [(H, NF), (1, H)]
    for _ in range(LAYERS):
        shapes += [(C, H), (H, G), (1, H)]
    shapes += [(NOUT, H), (1, NOUT)]
    return shapes


def param_count():
    return sum(r * c for r, c in layout())


def unpack(w):
    """Flat parameter vector -> dict of float64 arrays."""
    w = np.asarray(w, dtype=np.float64)
    out, off = [], 0
    for r, c in layout():
        a = w[off:off + r * c].reshape(r, c)
        out.append(a[0] if r == 1 else a)
        off += r * c
    p = {"We": out[0], "be": out[1], "Wd": out[-2], "bd": out[-1]}
    for l in range(LAYERS):
        p[f"Wu{l}"], p[f"Wo{l}"], p[f"bo{l}"] = out[2 + 3 * l], out[3 + 3 * l], out[4 + 3 * l]
    return p


def pack(p):
    parts = [p["We"], p["be"]]
    for l in range(LAYERS):
        parts += [p[f"Wu{l}"], p[f"Wo{l}"], p[f"bo{l}"]]
    parts += [p["Wd"], p["bd"]]
    return np.concatenate([np.ravel(a) for a in parts])


def init_params(seed=0, scale=1.0):
    """He-style random parameters (for tests)."""
    rng = np.random.default_rng(seed)
    parts = []
    for r, c in layout():
        if r == 1:
            parts.append(0.1 * scale * rng.standard_normal(c))
        else:
            parts.append(scale * np.sqrt(2.0 / c) * rng.standard_normal((r, c)))
    return np.concatenate([np.ravel(a) for a in parts])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def spline_shape(q):
    """Cubic-spline shape f(q) and f'(q) (src/kernel.cpp:20-55 without the
    10/(7 pi h^2) factor)."""
    q = np.asarray(q, dtype=np.float64)
    f = np.where(q < 1.0, 1.0 - 1.5 * q * q * (1.0 - 0.5 * q), np.where(q < 2.0, 0.25 * (2.0 - q) ** 3, 0.0))
    df = np.where(q < 1.0, -3.0 * q * (1.0 - 0.75 * q), np.where(q < 2.0, -0.75 * (2.0 - q) ** 2, 0.0))
    return f, df


def radius_graph(x, y, h):
    """Brute-force O(n^2) graph; rows sorted by neighbour index.
    Returns (row_ptr, col, edge3 [E x 3], wsum [n])."""
    x, y, h = (np.asarray(a, dtype=np.float64) for a in (x, y, h))
    n = len(x)
    rows, cols, e3 = [], [], []
    wsum = np.full(n, SPLINE_NORM)
    for i in range(n):
        dx = x[i] - x
        dy = y[i] - y
        r2 = dx * dx + dy * dy
        hij = 0.5 * (h[i] + h)
        m = r2 < 4.0 * hij * hij
        m[i] = False
        j = np.nonzero(m)[0]
        r = np.sqrt(r2[j])
        f, df = spline_shape(r / hij[j])
        w = SPLINE_NORM * f
        with np.errstate(invalid="ignore", divide="ignore"):
            sc = np.where(r > 0.0, SPLINE_NORM * df / r, 0.0)
        rows.append(np.full(len(j), i))
        cols.append(j)
        e3.append(np.stack([w, sc * dx[j], sc * dy[j]], axis=1))
        wsum[i] += w.sum()
    row = np.concatenate(rows) if rows else np.zeros(0, int)
    col = np.concatenate(cols) if cols else np.zeros(0, int)
    edge3 = np.concatenate(e3) if e3 else np.zeros((0, 3))
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_ptr, row + 1, 1)
    return np.cumsum(row_ptr), col, edge3, wsum


def _edge_rows(row_ptr):
    return np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))


def aggregate(row_ptr, col, edge3, U):
    """A[i] = [sum_e w U[j], sum_e gx U[j], sum_e gy U[j]] over the edges of i."""
    rows = _edge_rows(row_ptr)
    n = len(row_ptr) - 1
    A = np.zeros((n, 3 * U.shape[1]))
    Uj = U[col]
    for k in range(3):
        np.add.at(A[:, k * U.shape[1]:(k + 1) * U.shape[1]], rows, edge3[:, k:k + 1] * Uj)
    return A


def aggregate_T(row_ptr, col, edge3, dA):
    """Exact transpose of aggregate: dU[j] += sum_k e_k dA_k[i] for every edge i->j."""
    rows = _edge_rows(row_ptr)
    c = dA.shape[1] // 3
    dU = np.zeros((len(row_ptr) - 1, c))
    contrib = sum(edge3[:, k:k + 1] * dA[rows, k * c:(k + 1) * c] for k in range(3))
    np.add.at(dU, col, contrib)
    return dU


def forward(row_ptr, col, edge3, X, w):
    """X: n x 16 (normalised features).  Returns (Y [n x 8], cache)."""
    p = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    Z0 = X @ p["We"].T + p["be"]
    Hs = [silu(Z0)]
    Zs, Gs = [Z0], []
    for l in range(LAYERS):
        Hl = Hs[-1]
        U = Hl @ p[f"Wu{l}"].T
        A = aggregate(row_ptr, col, edge3, U)
        Gl = np.concatenate([Hl, A], axis=1)
        Z = Gl @ p[f"Wo{l}"].T + p[f"bo{l}"]
        Gs.append(Gl)
        Zs.append(Z)
        Hs.append(Hl + silu(Z))
    Y = Hs[-1] @ p["Wd"].T + p["bd"]
    return Y, dict(p=p, X=X, Z=Zs, H=Hs, G=Gs)


def loss_and_grad(row_ptr, col, edge3, X, T, w):
    """Mean-squared loss over the n x 8 outputs and its gradient (flat)."""
    Y, cc = forward(row_ptr, col, edge3, X, w)
    p = cc["p"]
    T = np.asarray(T, dtype=np.float64)
    n = Y.shape[0]
    R = Y - T
    loss = float((R * R).sum() / (n * NOUT))
    dY = 2.0 * R / (n * NOUT)
    g = {"Wd": dY.T @ cc["H"][-1], "bd": dY.sum(0)}
    dH = dY @ p["Wd"]
    for l in reversed(range(LAYERS)):
        dZ = dH * dsilu(cc["Z"][l + 1])
        Gl = cc["G"][l]
        g[f"Wo{l}"] = dZ.T @ Gl
        g[f"bo{l}"] = dZ.sum(0)
        dG = dZ @ p[f"Wo{l}"]
        dH = dH + dG[:, :H]
        dU = aggregate_T(row_ptr, col, edge3, dG[:, H:])
        g[f"Wu{l}"] = dU.T @ cc["H"][l]
        dH = dH + dU @ p[f"Wu{l}"]
    dZ0 = dH * dsilu(cc["Z"][0])
    g["We"] = dZ0.T @ cc["X"]
    g["be"] = dZ0.sum(0)
    return Y, loss, pack(g)


# ---- learned step (features and base update) -------------------------------

OUT_OF_FIELD = [0, 1, 2, 3, 7, 4, 5, 6]  # state field (x,y,vx,vy,rho,Sxx,Sxy,Syy) -> output index


def rk_stages(state8, deriv, tau, order=4):
    """One explicit RK step over tau of y' = deriv(y); returns (new state,
    [k1, ..., k_order]).  order 4: the classical RK4; order 3: Kutta's
    method (k2 at y + tau/2 k1, k3 at y - tau k1 + 2 tau k2,
    y + tau/6 (k1 + 4 k2 + k3))."""
    y = state8
    k1 = deriv(y)
    k2 = deriv(y + 0.5 * tau * k1)
    if order == 3:
        k3 = deriv(y - tau * k1 + 2.0 * tau * k2)
        return y + tau / 6.0 * (k1 + 4.0 * k2 + k3), [k1, k2, k3]
    k3 = deriv(y + 0.5 * tau * k2)
    k4 = deriv(y + tau * k3)
    return y + tau / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4), [k1, k2, k3, k4]


def stage_differences(ks, tau):
    """(Dlo, Dhi) from the stage rates, a_i = tau k_i.  Kutta (3 rates):
    D2 = 2 (a2 - a1), D3 = a3 - 2 a2 + a1; RK4 (4 rates): D3 = 4 (a3 - a2),
    D4 = 4 (a4 - a1 - D2 - D3/2).  For y' = A y, Dp = (tau A)^p y."""
    a = [tau * k for k in ks]
    d2 = 2.0 * (a[1] - a[0])
    if len(ks) == 3:
        return d2, a[2] - 2.0 * a[1] + a[0]
    d3 = 4.0 * (a[2] - a[1])
    return d3, 4.0 * (a[3] - a[0] - d2 - 0.5 * d3)


def features(start8, ks, vprev, h, wsum, tau, c0, K, rho0):
    """16 raw node features (same order as the library's node_features);
    ks = the base step's stage rates."""
    s = start8
    dlo, dhi = stage_differences(ks, tau)
    f = np.zeros((s.shape[0], NF))
    f[:, 0] = (s[:, 2] - vprev[:, 0]) / c0
    f[:, 1] = (s[:, 3] - vprev[:, 1]) / c0
    f[:, 2] = s[:, 5] / K
    f[:, 3] = s[:, 6] / K
    f[:, 4] = s[:, 7] / K
    f[:, 5] = s[:, 4] / rho0 - 1.0
    f[:, 6] = dlo[:, 2] / c0
    f[:, 7] = dlo[:, 3] / c0
    f[:, 8] = dhi[:, 2] / c0
    f[:, 9] = dhi[:, 3] / c0
    f[:, 10] = dhi[:, 5] / K
    f[:, 11] = dhi[:, 6] / K
    f[:, 12] = dhi[:, 7] / K
    f[:, 13] = dhi[:, 4] / rho0
    f[:, 14] = c0 * tau / h
    f[:, 15] = wsum
    return f


def _field_scales(h, c0, K, rho0):
    one = np.ones_like(h)
    return np.stack([h, h, c0 * one, c0 * one, rho0 * one, K * one, K * one, K * one], axis=1)


def targets(delta, h, c0, K, rho0):
    """Training targets: field increments (truth - base update, state-field
    order) made dimensionless, in network-output order (the library's
    k_targets)."""
    t = np.zeros_like(delta)
    t[:, OUT_OF_FIELD] = delta / _field_scales(h, c0, K, rho0)
    return t


def correction(Y, norm, h, c0, K, rho0):
    """Network outputs (normalised) -> increments of the 8 state fields."""
    tm = norm[2 * NF:2 * NF + NOUT]
    ts = norm[2 * NF + NOUT:]
    r = np.asarray(Y, dtype=np.float64) * ts + tm
    return r[:, OUT_OF_FIELD] * _field_scales(h, c0, K, rho0)


def conserve(delta, m):
    """Per-body projection of the increments of x, y, vx, vy onto zero
    mass-weighted mean (one body; the library does it per body)."""
    d = np.array(delta, dtype=np.float64, copy=True)
    M = m.sum()
    for k in range(4):
        d[:, k] -= (m * d[:, k]).sum() / M
    return d
