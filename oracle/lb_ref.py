"""CPU restatement of the learned thin boundary (test infrastructure only).

FP64 numpy forward/backward of the per-face network of include/fwe.h
(81 -> 64 -> 64 -> 9, SiLU, W1/W2/W3 used rounded to bfloat16), the 9-cell
stencil features of a layer cell, and the correction step (the dissipative
part of each correction, `shrink`): the checker for
libfwe.so's fwe_lb_* kernels.  The backward pass is pinned by central
finite differences (tests/test_lb_ref.py); there is no counterpart in the
reference, whose boundary is plain damping (OK4.m:297-341, ANIS.m:466-512).
"""
import numpy as np

NIN, NH, NOUT = 81, 64, 9


def nparams(L):
    return NIN * NH + L * NH + NH * NH + NH + NH * NOUT + NOUT


def unpack(p, L):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("w1", (NIN, NH)), ("b1", (L, NH)), ("w2", (NH, NH)), ("b2", (NH,)), ("w3", (NH, NOUT)),
                        ("b3", (NOUT,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == p.size
    return out


def pack(d):
    return np.concatenate([d[k].ravel() for k in ("w1", "b1", "w2", "b2", "w3", "b3")])


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(finite, r, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _sig(a):
    return 1.0 / (1.0 + np.exp(-a))


def forward(p, L, x, depth, rounded=True, act_bf16=False):
    """y [n,9] and the cache for backward.  rounded: weights used as
    bfloat16 (training and inference); act_bf16: inputs and hidden
    activations rounded to bfloat16 before each product (the matrix-core
    inference kernel)."""
    w = unpack(p, L)
    W1, W2, W3 = (bf16(w[k]) if rounded else w[k] for k in ("w1", "w2", "w3"))
    q = bf16 if act_bf16 else (lambda a: a)
    x = np.asarray(x, np.float64)
    a1 = q(x) @ W1 + w["b1"][np.asarray(depth)]
    h1 = a1 * _sig(a1)
    a2 = q(h1) @ W2 + w["b2"]
    h2 = a2 * _sig(a2)
    y = q(h2) @ W3 + w["b3"]
    return y, (x, np.asarray(depth), a1, h1, a2, h2, W2, W3)


def _dsilu(a):
    s = _sig(a)
    return s * (1 + a * (1 - s))


def loss_grad(p, L, x, depth, t, m, rounded=True):
    """loss = sum m (y - t)^2 / (9 n) and its gradient (packed layout)."""
    y, (x, d, a1, h1, a2, h2, W2, W3) = forward(p, L, x, depth, rounded)
    n = x.shape[0]
    e = y - t
    loss = float(np.sum(m * e * e) / (9 * n))
    g = 2 * m * e / (9 * n)
    da2 = (g @ W3.T) * _dsilu(a2)
    da1 = (da2 @ W2.T) * _dsilu(a1)
    grad = {"w3": h2.T @ g, "b3": g.sum(0), "w2": h1.T @ da2, "b2": da2.sum(0), "w1": x.T @ da1,
            "b1": np.zeros((L, NH))}
    np.add.at(grad["b1"], d, da1)
    return y, loss, pack(grad)


# ---------------------------------------------------------------------------
# layer cells and stencil features (fields in reference shapes, x fastest)
# ---------------------------------------------------------------------------
def layer_cells(nx, ny, nz, L):
    """[(i, j, k, face, depth)] grouped by face (x-, x+, y-, y+, z-, z+; the
    first face at the smallest distance), x fastest within a face."""
    per = [[] for _ in range(6)]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                dd = (i, nx - 1 - i, j, ny - 1 - j, k, nz - 1 - k)
                f = int(np.argmin(dd))
                if dd[f] < L:
                    per[f].append((i, j, k, f, dd[f]))
    return [c for f in range(6) for c in per[f]]


_NBR = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def stencil(face):
    ax = face // 2
    extra = []
    for s in (-2, 2):
        v = [0, 0, 0]
        v[ax] = s
        extra.append(tuple(v))
    return _NBR + extra


def features(fields, cells, nx, ny, nz, zinv):
    """x [n,81] (normalised), r [n], depth [n] of the cells; fields = list of
    the 9 arrays (Vx..Syz) in their staggered shapes."""
    n = len(cells)
    X = np.zeros((n, NIN))
    R = np.zeros(n)
    dep = np.zeros(n, np.int32)
    for s, (i, j, k, f, d) in enumerate(cells):
        v = np.zeros(NIN)
        for q, (di, dj, dk) in enumerate(stencil(f)):
            a, b, c = i + di, j + dj, k + dk
            if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                for comp in range(9):
                    val = fields[comp][a, b, c]
                    v[q * 9 + comp] = val * zinv if comp >= 3 else val
        r = np.sqrt(np.sum(v * v) / NIN) + 1e-30
        X[s] = v / r
        R[s] = r
        dep[s] = d
    return X, R, dep


def live(i, j, k, comp):
    if comp == 6:
        return i > 0 and j > 0
    if comp == 7:
        return i > 0 and k > 0
    if comp == 8:
        return j > 0 and k > 0
    return True


def shrink(v, d):
    """Dissipative part of correction d applied to the stepped value v: v
    moves toward zero by at most |v| and never grows."""
    if v > 0:
        return v + min(max(d, -v), 0.0)
    if v < 0:
        return v + max(min(d, -v), 0.0)
    return v


def corrections(params, L, fields, nx, ny, nz, z):
    """Per-cell corrections [(i, j, k, comp, value)] of the layer, from the
    fields before a step, as the inference kernel computes them (bfloat16
    inputs and activations); params = list of 6 packed parameter vectors."""
    cells = layer_cells(nx, ny, nz, L)
    X, R, dep = features(fields, cells, nx, ny, nz, 1.0 / z)
    out = []
    faces = np.array([c[3] for c in cells])
    for f in range(6):
        sel = np.nonzero(faces == f)[0]
        if sel.size == 0:
            continue
        y, _ = forward(params[f], L, X[sel], dep[sel], act_bf16=True)
        for row, s in enumerate(sel):
            i, j, k = cells[s][:3]
            for comp in range(9):
                if live(i, j, k, comp):
                    out.append((i, j, k, comp, y[row, comp] * R[s] * (z if comp >= 3 else 1.0)))
    return out
