"""Host construction of the shifted-Laplacian multigrid hierarchy.

Level l+1 keeps the even nodes of level l plus its last node on every axis,
so any grid size coarsens (the last coarse interval may be one fine cell).
Each level is the finite-volume operator rediscretised on its node set:

* per-axis dual widths and edge lengths in units of the fine spacing h
  (integrated rows keep the fine h^3 normalisation on every level, which
  makes P^T the consistent restriction);
* PML stretch sampled at the coarse nodes (sigma is a function of position);
* 1/rho and kappa = 1/(rho c^2) restricted by normalised full weighting;
* open faces get i omega area sqrt(kappa/rho), impedance faces i omega area / Z.

By default (device_fields) the library derives the 3-D fields and the open
faces' terms on the device from the operator's own fields; the host builds
only the O(n) axis tables and the O(n^2) impedance terms.  The host
restriction below is the same definition, kept for checking the device one.

The library shifts every level's mass term by omega^2 -> omega^2 (1 + i beta).
None of this exists in the reference (its only preconditioner is Jacobi,
src/algorithm/solvers/iterative_solver.py:46-56); it is an opt-in
performance mode for the mixed form, checked by solution parity.
"""
import ctypes

import numpy as np

from . import _capi, physics


def coarse_nodes(n):
    idx = np.arange(0, n, 2)
    if idx[-1] != n - 1:
        idx = np.append(idx, n - 1)
    return idx


def prolongation_1d(n):
    """Dense (n, m) 1-D trilinear prolongation for the even-node coarsening."""
    c = coarse_nodes(n)
    where = {int(f): i for i, f in enumerate(c)}
    p = np.zeros((n, len(c)))
    for f in range(n):
        if f in where:
            p[f, where[f]] = 1.0
        else:
            p[f, where[f - 1]] = 0.5
            p[f, where[f + 1]] = 0.5
    return p


def _restrict_field(field, mats):
    """Normalised full weighting of a nodal field: (P^T f) / (P^T 1), axis by axis."""
    out = field
    for axis, p in enumerate(mats):
        pt = p.T/p.sum(axis=0)[:, None]
        out = np.moveaxis(np.tensordot(pt, np.moveaxis(out, axis, 0), axes=(1, 0)), 0, axis)
    return out


def _axis_tables(x):
    """Dual widths and edge lengths (units of h) of node positions x (units of h)."""
    e = np.diff(x)
    w = np.empty(len(x))
    w[1:-1] = 0.5*(x[2:] - x[:-2])
    w[0] = 0.5*e[0]
    w[-1] = 0.5*e[-1]
    return w, e


# Nodes of the largest level the library solves in LDS, one workgroup per
# column (its coarsest-level solver); coarsening stops there by default.
COARSE_MAX_NODES = 1024


class Hierarchy:
    def __init__(self, op, levels=None, min_size=5, device_fields=True):
        self.levels = []
        self.device_fields = bool(device_fields)
        shape = op.shape
        x = [np.arange(n, dtype=float) for n in shape]
        sigma = [np.imag(s) for s in op.stretch]
        inv_rho = kappa = None
        if not self.device_fields:
            inv_rho = 1.0/op.rho
            kappa = 1.0/(op.rho*op.c**2)
        kinds = op.kinds
        while True:
            lev = {"shape": shape, "inv_rho": inv_rho, "kappa": kappa}
            tabs = [_axis_tables(x[a]) for a in range(3)]
            lev["weight"] = [t[0] for t in tabs]
            lev["edge"] = [t[1] for t in tabs]
            lev["stretch"] = [1.0 + 1j*sigma[a] for a in range(3)]
            if not self.levels:
                lev["faces"] = op.face_terms
            else:
                lev["faces"] = self._faces(op, lev, kinds)
            self.levels.append(lev)
            if levels is not None and len(self.levels) == levels:
                break
            if min(shape) <= min_size or (levels is None and (len(self.levels) >= 8
                                                               or int(np.prod(shape)) <= COARSE_MAX_NODES)):
                break
            idx = [coarse_nodes(n) for n in shape]
            if not self.device_fields:
                mats = [prolongation_1d(n) for n in shape]
                inv_rho = _restrict_field(inv_rho, mats)
                kappa = _restrict_field(kappa, mats)
            x = [x[a][idx[a]] for a in range(3)]
            sigma = [sigma[a][idx[a]] for a in range(3)]
            shape = tuple(len(i) for i in idx)

    def _faces(self, op, lev, kinds):
        out = []
        for index, kind in enumerate(kinds):
            if kind not in ("open", "impedance"):
                out.append(None)
                continue
            axis = index//2
            b, c = [a for a in range(3) if a != axis]
            w, s = lev["weight"], lev["stretch"]
            area = (w[b][:, None]*w[c][None, :]*s[b][:, None]*s[c][None, :])/op.dx
            sel = [slice(None)]*3
            sel[axis] = 0 if index % 2 == 0 else -1
            sel = tuple(sel)
            if kind == "open":
                if self.device_fields:
                    out.append(None)  # derived on the device
                    continue
                adm = np.sqrt(lev["kappa"][sel]*lev["inv_rho"][sel])
            else:
                z = np.asarray(op.boundaries[physics.FACES[index]]["impedance"], float)
                adm = 1.0/complex(z[0], z[1])
            out.append(np.ascontiguousarray((1j*op.omega*area*adm).ravel(order="F"), dtype=np.complex128))
        return out

    def install(self, op, shift=1.2, pre=1, post=1, coarse_sweeps=30, weight=0.7, cycle=1, post_coarse=5):
        keep = []

        def ptr(a):
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        descs = (_capi.LevelDesc*len(self.levels))()
        null = ctypes.POINTER(ctypes.c_double)()
        for d, lev in zip(descs, self.levels):
            d.nx, d.ny, d.nz = lev["shape"]
            if self.device_fields:
                d.rho = d.kappa = null
            else:
                d.rho = ptr((1.0/lev["inv_rho"]).ravel(order="F"))
                d.kappa = ptr(lev["kappa"].ravel(order="F"))
            for a in range(3):
                d.weight[a] = ptr(lev["weight"][a])
                d.stretch[a] = ptr(lev["stretch"][a].astype(np.complex128).view(np.float64))
                d.edge[a] = ptr(lev["edge"][a])
            for f in range(6):
                t = lev["faces"][f]
                d.face_term[f] = ptr(t.view(np.float64)) if t is not None else ctypes.POINTER(ctypes.c_double)()
        _capi.check(_capi.lib().hz_set_multigrid(op.handle, len(self.levels), descs, float(shift), int(pre),
                                                 int(post), int(post_coarse), int(coarse_sweeps), float(weight),
                                                 int(cycle)),
                    "hz_set_multigrid")
        return self
