"""Independent NumPy restatement of the reference multigrid (mg.wgsl, sim.rs:451-491).

Vectorised FP64 array code, written separately from oracle/pyro_oracle.c, used
only to cross-check the oracle's loop-based V-cycle (indexing of restriction,
prolongation and the red-black order).
"""
import numpy as np


def nb_sum(p, f):
    """Sum of fluid-neighbour phi and the count cnt of mg.wgsl:25-33."""
    fl = f == 0
    q = np.where(fl, p, 0.0)
    s = np.zeros_like(p)
    c = np.zeros_like(p)
    s[:, :, :-1] += q[:, :, 1:]; c[:, :, :-1] += fl[:, :, 1:]; c[:, :, -1] += 1
    s[:, :, 1:] += q[:, :, :-1]; c[:, :, 1:] += fl[:, :, :-1]; c[:, :, 0] += 1
    s[:, :-1, :] += q[:, 1:, :]; c[:, :-1, :] += fl[:, 1:, :]; c[:, -1, :] += 1
    s[:, 1:, :] += q[:, :-1, :]; c[:, 1:, :] += fl[:, :-1, :]  # ground: not counted
    s[:-1] += q[1:]; c[:-1] += fl[1:]; c[-1] += 1
    s[1:] += q[:-1]; c[1:] += fl[:-1]; c[0] += 1
    return s, c


def resid(p, b, f, h):
    s, c = nb_sum(p, f)
    return np.where(f == 0, b - (s - c * p) / (h * h), 0.0)


def smooth(p, b, f, h, iters, om):
    nz, ny, nx = p.shape
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    col = (X + Y + Z) & 1
    for _ in range(iters):
        for cc in (0, 1):
            s, c = nb_sum(p, f)
            new = p + ((s - h * h * b) / np.maximum(c, 1) - p) * om
            p = np.where((col == cc) & (f == 0) & (c > 0.5), new, p)
    return p


def restrict(p, b, f, h):
    r = resid(p, b, f, h)
    nz, ny, nx = p.shape
    cz, cy, cx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    rs = np.zeros((cz, cy, cx))
    n = np.zeros((cz, cy, cx))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                sub = r[dz::2, dy::2, dx::2]
                fl = f[dz::2, dy::2, dx::2] == 0
                sl = (slice(0, sub.shape[0]), slice(0, sub.shape[1]), slice(0, sub.shape[2]))
                rs[sl] += sub
                n[sl] += fl
    return np.where(n > 0, rs / np.maximum(n, 1), 0.0), np.where(n > 0, 0, 1).astype(np.uint32)


def prolong(pc, fc, pf, ff):
    nz, ny, nx = pf.shape
    cz, cy, cx = pc.shape

    def ax(n):
        x = (np.arange(n) + 0.5) * 0.5 - 0.5
        i0 = np.floor(x).astype(int)
        return i0, x - i0

    iz, fz = ax(nz)
    iy, fy = ax(ny)
    ix, fx = ax(nx)
    acc = np.zeros_like(pf)
    ws = np.zeros_like(pf)
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                zz, yy, xx = np.clip(iz + oz, 0, cz - 1), np.clip(iy + oy, 0, cy - 1), np.clip(ix + ox, 0, cx - 1)
                w = (fz if oz else 1 - fz)[:, None, None] * (fy if oy else 1 - fy)[None, :, None] * \
                    (fx if ox else 1 - fx)[None, None, :]
                m = fc[np.ix_(zz, yy, xx)] == 0
                acc += np.where(m, w * pc[np.ix_(zz, yy, xx)], 0.0)
                ws += np.where(m, w, 0.0)
    return np.where((ff == 0) & (ws > 1e-4), pf + acc / np.maximum(ws, 1e-30), pf)


def vcycle(p, b, f, h, nl, coarse_iters=80):
    ps, bs, fs, hs = [p], [b], [f], [h]
    for l in range(nl - 1):
        ps[l] = smooth(ps[l], bs[l], fs[l], hs[l], 2, 1.0)
        bc, fc = restrict(ps[l], bs[l], fs[l], hs[l])
        ps.append(np.zeros_like(bc)); bs.append(bc); fs.append(fc); hs.append(hs[l] * 2)
    ps[-1] = smooth(ps[-1], bs[-1], fs[-1], hs[-1], coarse_iters, 1.6)
    for l in range(nl - 2, -1, -1):
        ps[l] = prolong(ps[l + 1], fs[l + 1], ps[l], fs[l])
        ps[l] = smooth(ps[l], bs[l], fs[l], hs[l], 2, 1.0)
    return ps[0]
