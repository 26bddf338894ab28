"""Tile-major <-> global array helpers shared by the tests."""
import numpy as np


def gather(buf, ntiles, nx, ext_x, vol, off, ncomp=3):
    """Interior cells of a tile-major buffer as a global [ny_total, nx_total, ncomp] array."""
    ntx, nty = ntiles
    b = np.asarray(buf).reshape(nty, ntx, vol, ncomp)
    rows = b[:, :, off: off + nx[1] * ext_x].reshape(nty, ntx, nx[1], ext_x, ncomp)[:, :, :, :nx[0]]
    out = rows.transpose(0, 2, 1, 3, 4).reshape(nty * nx[1], ntx * nx[0], ncomp)
    return out if ncomp > 1 else out[..., 0]


def scatter_full(fn, ntiles, nx, gl, gu, dtype=np.float64, ncomp=3):
    """Tile-major buffer whose every element (guards included) holds fn(gx, gy)
    evaluated at the element's global (unwrapped) cell index."""
    ntx, nty = ntiles
    ext_x, ext_y = gl + nx[0] + gu, gl + nx[1] + gu
    vol = (ext_x * ext_y + 3) & ~3
    out = np.zeros((ntx * nty, vol, ncomp), dtype)
    for t in range(ntx * nty):
        tx, ty = t % ntx, t // ntx
        for j in range(ext_y):
            for i in range(ext_x):
                gx = tx * nx[0] + i - gl
                gy = ty * nx[1] + j - gl
                out[t, j * ext_x + i] = fn(gx, gy)
    return out.reshape(-1, ncomp)


def global_positions(p, ntiles, nx):
    """Global cell index (x, y) of every particle from its tile and tile-local cell."""
    ntx = ntiles[0]
    tile = np.repeat(np.arange(len(p["np"])), p["np"])
    gx = (tile % ntx) * nx[0] + p["ix"][:, 0]
    gy = (tile // ntx) * nx[1] + p["ix"][:, 1]
    return gx, gy
