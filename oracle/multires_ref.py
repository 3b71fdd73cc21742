"""CPU oracle of the per-block-pitch field.  TEST INFRASTRUCTURE: only the
tests use it, as the checker of the package's multires module.

Restates FE/faltkarna_v1_multires.py with numpy and scipy.ndimage (the
reference's own third-party calls): level balancing :171-190 (grey_dilation
over a 3x3x3 cube, then 2:1 over the face cross), the overlap band :211-252
(map_coordinates order 1, mode "nearest" on the coarser neighbour's tile, faces
in axis order, lower side first, level 1 bricks before level 2), the brick of
a point :386-402, avstand :405-441, gradient :444-460 and kontinuitetscheck
:462-520.  Tile samples are the fp32 recipe fold of oracle/fe_oracle.c
(or_recipe_eval_points) at the positions the device uses: level 0 at
x0 + f32(min(i, n-1)) * h, level l at x0 + f32(i) * f32(h / 2^l).  The
per-brick level map is an input, as in the package.
"""
import math

import numpy as np
from scipy import ndimage

import oracle as O

NIVAER = (0, 1, 2)
F32 = np.float32


def balance(niva, aktiv):
    n = niva.copy().astype(np.int8)
    n[~aktiv] = -1
    grown = ndimage.grey_dilation(np.where(aktiv, n, -1), size=(3, 3, 3), mode="nearest")
    n = np.where(aktiv, np.maximum(n, grown), -1).astype(np.int8)
    for _ in range(len(NIVAER)):
        omg = ndimage.grey_dilation(np.where(aktiv, n, -1), footprint=ndimage.generate_binary_structure(3, 1),
                                    mode="nearest")
        ny = np.where(aktiv, np.maximum(n, omg - 1), -1).astype(np.int8)
        if np.array_equal(ny, n):
            break
        n = ny
    return n


class Field:
    def __init__(self, x0, y0, z0, pitch, nx, ny, nz, block, recipe, niva, band=True):
        self.x0, self.y0, self.z0 = F32(x0), F32(y0), F32(z0)
        self.pitch = F32(pitch)
        self.n = (nx, ny, nz)
        self.block = block
        B = block
        self.nb = tuple(-(-v // B) for v in self.n)
        g = O.grid_record(self.x0, self.y0, self.z0, self.pitch, nx, ny, nz, B, F32(math.sqrt(3.0) * B * pitch))
        prims = np.ascontiguousarray(recipe).view(O.PRIM)
        kind, _cma, _ids, _tiles = O.build_sparse(prims, g)
        self.kind = kind.reshape(self.nb).astype(np.int8)
        aktiv = self.kind == 0
        self.niva = balance(np.where(aktiv, niva, -1).astype(np.int8), aktiv)
        self.lo = np.array([float(self.x0), float(self.y0), float(self.z0)])
        self.h = float(self.pitch)
        self.L = B * self.h
        self.margin = math.sqrt(3.0) * B * self.h
        self.tiles = {}
        for bi, bj, bk in np.argwhere(aktiv):
            self.tiles[(int(bi) * self.nb[1] + int(bj)) * self.nb[2] + int(bk)] = self._tile(prims, bi, bj, bk)
        if band:
            for lvl in NIVAER[1:]:
                for lin in sorted(self.tiles):
                    if self.level_of(lin) == lvl:
                        self._band(lin)

    def brick_of_lin(self, lin):
        nby, nbz = self.nb[1], self.nb[2]
        return lin // (nby * nbz), (lin // nbz) % nby, lin % nbz

    def level_of(self, lin):
        return int(self.niva[self.brick_of_lin(lin)])

    def pitch_of(self, lvl):
        return self.h / float(1 << int(lvl))

    def _tile(self, prims, bi, bj, bk):
        lvl = int(self.niva[bi, bj, bk])
        n = self.block << lvl
        m = n + 1
        idx = np.arange(m)
        org = (self.x0, self.y0, self.z0)
        coords = []
        for a, b in enumerate((bi, bj, bk)):
            if lvl == 0:
                gi = np.minimum(b * self.block + idx, self.n[a] - 1)
                coords.append((org[a] + gi.astype(F32) * self.pitch).astype(F32))
            else:
                hl = F32(np.ldexp(self.pitch, -lvl))
                coords.append((org[a] + (b * n + idx).astype(F32) * hl).astype(F32))
        X, Y, Z = np.meshgrid(*coords, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        return O.recipe_eval_points(prims, pts).reshape(m, m, m)

    def _band(self, lin):
        bi, bj, bk = self.brick_of_lin(lin)
        lvl = self.level_of(lin)
        n_l = self.block * (1 << lvl)
        tile = self.tiles[lin]
        b_origin = self.lo + np.array([bi, bj, bk], dtype=np.float64) * self.L
        for ax in range(3):
            for side in (-1, +1):
                nb = [bi, bj, bk]
                nb[ax] += side
                if not all(0 <= nb[a] < self.nb[a] for a in range(3)) or self.kind[tuple(nb)] != 0:
                    continue
                lvl_n = int(self.niva[tuple(nb)])
                if lvl_n >= lvl:
                    continue
                nlin = (nb[0] * self.nb[1] + nb[1]) * self.nb[2] + nb[2]
                idx = [np.arange(n_l + 1, dtype=np.float64)] * 3
                idx[ax] = np.array([0.0 if side < 0 else float(n_l)])
                gi, gj, gk = np.meshgrid(*idx, indexing="ij")
                pts = np.stack([gi.ravel(), gj.ravel(), gk.ravel()], axis=1) * self.pitch_of(lvl) + b_origin
                nb_origin = self.lo + np.array(nb, dtype=np.float64) * self.L
                lok = (pts - nb_origin) / self.pitch_of(lvl_n)
                v = ndimage.map_coordinates(self.tiles[nlin], lok.T, order=1, mode="nearest")
                sl = [slice(None)] * 3
                sl[ax] = 0 if side < 0 else n_l
                tile[tuple(sl)] = v.reshape(tile[tuple(sl)].shape).astype(np.float32)

    def bricks_of(self, pts):
        rel = (np.asarray(pts, dtype=np.float64) - self.lo) / self.L
        b = np.floor(rel).astype(np.int64)
        shp = np.array(self.nb, dtype=np.int64)
        outside = np.any((b < 0) | (b >= shp), axis=1)
        b = np.clip(b, 0, shp - 1)
        inactive = self.kind[b[:, 0], b[:, 1], b[:, 2]] != 0
        for ax in range(3):
            on = (rel[:, ax] == np.floor(rel[:, ax])) & (b[:, ax] > 0) & inactive & ~outside
            if not on.any():
                continue
            cand = b[on].copy()
            cand[:, ax] -= 1
            ok = self.kind[cand[:, 0], cand[:, 1], cand[:, 2]] == 0
            idx = np.where(on)[0][ok]
            b[idx, ax] -= 1
            inactive = self.kind[b[:, 0], b[:, 1], b[:, 2]] != 0
        return b, outside

    def avstand(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        b, outside = self.bricks_of(pts)
        lin = (b[:, 0] * self.nb[1] + b[:, 1]) * self.nb[2] + b[:, 2]
        kd = self.kind[b[:, 0], b[:, 1], b[:, 2]]
        out = np.where(kd < 0, -self.margin, self.margin).astype(np.float64)
        out[outside] = self.margin
        act = (kd == 0) & ~outside
        for u in np.unique(lin[act]):
            sel = act & (lin == u)
            bi, bj, bk = self.brick_of_lin(int(u))
            o = self.lo + np.array([bi, bj, bk], dtype=np.float64) * self.L
            lok = (pts[sel] - o) / self.pitch_of(self.niva[bi, bj, bk])
            out[sel] = ndimage.map_coordinates(self.tiles[int(u)], lok.T, order=1, mode="nearest")
        return out

    def gradient(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        b, _ = self.bricks_of(pts)
        lvl = self.niva[b[:, 0], b[:, 1], b[:, 2]].astype(np.int64)
        h = self.h / np.power(2.0, np.maximum(lvl, 0))
        o = self.lo + b.astype(np.float64) * self.L
        g = np.zeros((len(pts), 3))
        for ax in range(3):
            d = np.zeros((len(pts), 3))
            d[:, ax] = h
            pa = np.minimum(pts + 0.5 * d, o + self.L)
            pb = np.maximum(pts - 0.5 * d, o)
            g[:, ax] = (self.avstand(pa) - self.avstand(pb)) / np.maximum(pa[:, ax] - pb[:, ax], 1e-12)
        return g

    def continuity(self, eps_frac=1.0e-3, interior=False):
        """(faces, points, max jump, bound) over the pitch-change faces;
        interior: only the face lattice's interior points (the package's
        bara_inre extension)."""
        faces, npts, mx, bound = 0, 0, 0.0, 0.0
        for lin in sorted(self.tiles):
            bi, bj, bk = self.brick_of_lin(lin)
            lvl = self.level_of(lin)
            n_l = self.block * (1 << lvl)
            h_f = self.pitch_of(lvl)
            b_origin = self.lo + np.array([bi, bj, bk], dtype=np.float64) * self.L
            for ax in range(3):
                for side in (-1, +1):
                    nb = [bi, bj, bk]
                    nb[ax] += side
                    if not all(0 <= nb[a] < self.nb[a] for a in range(3)) or self.kind[tuple(nb)] != 0:
                        continue
                    lvl_n = int(self.niva[tuple(nb)])
                    if lvl_n >= lvl:
                        continue
                    idx = [np.arange(1 if interior else 0, n_l if interior else n_l + 1, dtype=np.float64)] * 3
                    idx[ax] = np.array([0.0 if side < 0 else float(n_l)])
                    gi, gj, gk = np.meshgrid(*idx, indexing="ij")
                    pts = np.stack([gi.ravel(), gj.ravel(), gk.ravel()], axis=1) * h_f + b_origin
                    e = np.zeros(3)
                    e[ax] = eps_frac * h_f
                    jump = np.abs(self.avstand(pts - side * e) - self.avstand(pts + side * e))
                    faces += 1
                    npts += len(pts)
                    mx = max(mx, float(jump.max()))
                    bound = max(bound, 0.5 * self.pitch_of(lvl_n))
        return faces, npts, mx, bound
