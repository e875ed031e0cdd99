"""Independent numpy finite-element model used to pin the oracle.

Shape-function derivatives are obtained by central differences of the
trilinear shape functions N_a(x) = prod_d (x_d if bit_d(a) else 1 - x_d), and
element matrices by 3x3x3 Gauss quadrature -- none of the oracle's closed-form
B-matrix code or its 2x2x2 rule is reused.  The global operator is assembled as
a scipy sparse matrix over the periodic node set.
"""
import numpy as np
import scipy.sparse as sp

R2 = np.sqrt(0.5)


def shape(a, xi):
    bits = [(a >> d) & 1 for d in range(3)]
    return np.prod([xi[d] if bits[d] else 1.0 - xi[d] for d in range(3)])


def dshape(xi, le, h=1e-6):
    """dN_a/dx_d at reference point xi (8 x 3), by central differences."""
    out = np.zeros((8, 3))
    for a in range(8):
        for d in range(3):
            p = np.array(xi, float)
            m = np.array(xi, float)
            p[d] += h
            m[d] -= h
            out[a, d] = (shape(a, p) - shape(a, m)) / (2 * h) / le[d]
    return out


def bmat(xi, le, howmany):
    """Strain-displacement matrix (Mandel for howmany=3), nodes a-major, comps c-minor."""
    dN = dshape(xi, le)
    if howmany == 1:
        return dN.T.copy()  # 3 x 8
    B = np.zeros((6, 24))
    for a in range(8):
        gx, gy, gz = dN[a]
        B[0, 3 * a + 0] = gx
        B[1, 3 * a + 1] = gy
        B[2, 3 * a + 2] = gz
        B[3, 3 * a + 0], B[3, 3 * a + 1] = R2 * gy, R2 * gx
        B[4, 3 * a + 0], B[4, 3 * a + 2] = R2 * gz, R2 * gx
        B[5, 3 * a + 1], B[5, 3 * a + 2] = R2 * gz, R2 * gy
    return B


def gauss_points(order):
    x, w = np.polynomial.legendre.leggauss(order)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    pts = []
    for k in range(order):
        for j in range(order):
            for i in range(order):
                pts.append(((x[i], x[j], x[k]), w[i] * w[j] * w[k]))
    return pts


def iso_C(howmany, p):
    if howmany == 1:
        return p[0] * np.eye(3)
    K, G = p
    lam = K - 2.0 / 3.0 * G
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C += 2 * G * np.eye(6)
    return C


def element_matrices(C, le, howmany, order=3):
    """Ke = int B^T C B dV and Fe = int B^T C dV (for a macroscopic gradient)."""
    ve = float(np.prod(le))
    ne = 8 * howmany
    Ke = np.zeros((ne, ne))
    Fe = np.zeros((ne, C.shape[0]))
    for xi, w in gauss_points(order):
        B = bmat(xi, le, howmany)
        Ke += w * ve * B.T @ C @ B
        Fe += w * ve * B.T @ C
    return Ke, Fe


def node_index(dims, x, y, z):
    nx, ny, nz = dims
    return ((x % nx) * ny + (y % ny)) * nz + (z % nz)


def element_dofs(dims, howmany, x, y, z):
    nodes = [node_index(dims, x + (a & 1), y + ((a >> 1) & 1), z + ((a >> 2) & 1)) for a in range(8)]
    return np.array([n * howmany + c for n in nodes for c in range(howmany)])


def assemble(ms, Cs, le, howmany):
    """Global sparse K and the macroscopic-gradient load operator F (ndof x n_str)."""
    dims = ms.shape
    N = int(np.prod(dims))
    ndof = N * howmany
    mats = {p: element_matrices(Cs[p], le, howmany) for p in np.unique(ms)}
    rows, cols, vals = [], [], []
    F = np.zeros((ndof, Cs[int(ms.flat[0])].shape[0]))
    for x in range(dims[0]):
        for y in range(dims[1]):
            for z in range(dims[2]):
                Ke, Fe = mats[int(ms[x, y, z])]
                dofs = element_dofs(dims, howmany, x, y, z)
                rows.append(np.repeat(dofs, len(dofs)))
                cols.append(np.tile(dofs, len(dofs)))
                vals.append(Ke.reshape(-1))
                np.add.at(F, dofs, Fe)
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(ndof, ndof))
    return K, F


def solve_periodic(ms, Cs, le, howmany, g0):
    """Direct solve of K u = -F g0 with zero-mean gauge; returns (u [N,h], mean stress)."""
    import scipy.sparse.linalg as spla
    K, F = assemble(ms, Cs, le, howmany)
    N = ms.size
    ndof = N * howmany
    # gauge: append mean-zero constraints with Lagrange multipliers
    Cg = sp.lil_matrix((howmany, ndof))
    for c in range(howmany):
        Cg[c, c::howmany] = 1.0
    A = sp.bmat([[K, Cg.T], [Cg, None]]).tocsc()
    rhs = np.concatenate([-F @ g0, np.zeros(howmany)])
    sol = spla.spsolve(A, rhs)
    u = sol[:ndof].reshape(ms.shape + (howmany,))
    # homogenized stress: volume mean of C (B u + g0)
    dims = ms.shape
    ns = len(g0)
    acc = np.zeros(ns)
    pts = gauss_points(2)
    for x in range(dims[0]):
        for y in range(dims[1]):
            for z in range(dims[2]):
                C = Cs[int(ms[x, y, z])]
                ue = sol[element_dofs(dims, howmany, x, y, z)]
                s = np.zeros(ns)
                for xi, w in pts:
                    s += w * C @ (bmat(xi, le, howmany) @ ue + g0)
                acc += s
    return u, acc / ms.size
