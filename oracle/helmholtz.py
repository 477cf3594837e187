"""FP64 restatement of the conservative finite-volume Helmholtz operator.

Follows ``src/algorithm/physics/conservative_helmholtz.py`` (reference):
per-axis half-volume weights (:48-53), PML stretch (:54-64), mass weight
(:65-66), soft/fixed rows (:85-88), open/impedance face terms (:89-103),
axis edges with the midpoint stretch (:105-117), CSR assembly (:119-149) and
the material VJP (:151-175).  Unknowns are numbered Fortran-style,
p = i + nx*(j + ny*k).

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
import scipy.sparse as sp

FACE_NAMES = ("-x", "+x", "-y", "+y", "-z", "+z")
FACE_KINDS = ("reflecting", "open", "soft", "impedance")


def face_kind(spec):
    """``boundary_type`` (conservative_helmholtz.py:13-14)."""
    if isinstance(spec, dict):
        return spec.get("type", "reflecting")
    return spec


class OracleOperator:
    """Dense description of A; ``matrix()`` returns the complex128 CSR."""

    def __init__(self, rho, c, dx, omega, boundaries, pml=None, fixed_dofs=()):
        rho = np.asarray(rho, dtype=np.float64)
        c = np.asarray(c, dtype=np.float64)
        if rho.ndim != 3 or c.shape != rho.shape or min(rho.shape) < 2:
            raise ValueError("rho and c must share a 3-D grid with every edge >= 2")
        if not (np.isfinite(rho).all() and np.isfinite(c).all() and (rho > 0).all() and (c > 0).all()):
            raise ValueError("density and sound speed must be finite and positive")
        if not (np.isfinite(dx) and dx > 0 and np.isfinite(omega) and omega > 0):
            raise ValueError("dx and omega must be finite and positive")
        unknown = set(boundaries) - set(FACE_NAMES)
        if unknown:
            raise ValueError("Unknown boundary face")
        self.rho, self.c = rho, c
        self.shape = rho.shape
        self.dx, self.omega = float(dx), float(omega)
        self.boundaries = dict(boundaries)
        pml = dict(pml or {})
        cells = int(pml.get("cells", 0))
        strength = float(pml.get("strength", 2.0))
        order = int(pml.get("order", 3))
        if cells < 0 or order < 1 or not np.isfinite(strength) or strength < 0:
            raise ValueError("Invalid PML cells, strength or order")
        default_faces = [f for f in FACE_NAMES
                         if face_kind(self.boundaries.get(f, "reflecting")) == "open"]
        pml_faces = list(pml.get("faces", default_faces))
        if set(pml_faces) - set(FACE_NAMES):
            raise ValueError("Unknown PML face")

        # 1-D tables per axis (conservative_helmholtz.py:48-64).
        self.w1d, self.s1d = [], []
        for axis, n in enumerate(self.shape):
            w = np.ones(n)
            w[0] = w[-1] = 0.5
            sigma = np.zeros(n)
            faces_here = [f for f in FACE_NAMES[2*axis:2*axis + 2] if f in pml_faces]
            if cells and faces_here:
                if cells*len(faces_here) >= n - 1:
                    raise ValueError("PML leaves no physical interior on an axis")
                for f in faces_here:
                    if face_kind(self.boundaries.get(f, "reflecting")) != "open":
                        raise ValueError("PML may only be placed on open faces")
                    d = np.arange(n, dtype=np.float64)
                    if f[0] == "+":
                        d = d[::-1]
                    sigma = sigma + strength*np.maximum(1.0 - d/cells, 0.0)**order
            self.w1d.append(w)
            self.s1d.append(1.0 + 1j*sigma)

        wx, wy, wz = self._grid(self.w1d)
        sx, sy, sz = self._grid(self.s1d)
        self.volume = wx*wy*wz
        self.mass_weight = self.volume*sx*sy*sz

        self.fixed = np.zeros(self.shape, dtype=bool)
        self.fixed.ravel(order="F")[np.asarray(fixed_dofs, dtype=int)] = True
        pml_nodes = (np.imag(sx) + np.imag(sy) + np.imag(sz)) > 0
        if np.any(self.fixed & pml_nodes):
            raise ValueError("PML overlaps a source; enlarge the domain or reduce PML thickness")
        source_fixed = self.fixed.copy()

        # Face terms (conservative_helmholtz.py:78-103); order of the loop is -x,+x,-y,+y,-z,+z.
        self.face_terms = []
        for index, name in enumerate(FACE_NAMES):
            axis = index//2
            sel = self._face_selection(axis, index % 2)
            spec = self.boundaries.get(name, "reflecting")
            kind = face_kind(spec)
            if kind not in FACE_KINDS:
                raise ValueError(f"Unsupported boundary {name}: {spec!r}")
            if kind == "soft":
                if np.any(source_fixed[sel]):
                    raise ValueError("A pressure-release face conflicts with a source")
                self.fixed[sel] = True
            elif kind in ("open", "impedance"):
                b, c_ = [a for a in range(3) if a != axis]
                tables = [self._axis_view(self.w1d[a], a) for a in range(3)]
                stretches = [self._axis_view(self.s1d[a], a) for a in range(3)]
                area = tables[b]*tables[c_]*stretches[b]*stretches[c_]
                area = np.broadcast_to(area, self.shape)[sel]/self.dx
                if kind == "open":
                    admittance = 1.0/(self.rho[sel]*self.c[sel])
                else:
                    if not isinstance(spec, dict):
                        raise ValueError("impedance requires {type: impedance, impedance: [real, imag]}")
                    z = np.asarray(spec.get("impedance"), dtype=float)
                    if z.shape != (2,) or not np.isfinite(z).all() or z[0] < 0 or not np.any(z):
                        raise ValueError("impedance must be finite [real>=0, imag], nonzero, in Pa s/m")
                    admittance = 1.0/complex(z[0], z[1])
                self.face_terms.append((sel, kind, 1j*self.omega*area*admittance))

    # -- helpers ---------------------------------------------------------
    def _axis_view(self, table, axis):
        dims = [1, 1, 1]
        dims[axis] = len(table)
        return table.reshape(dims)

    def _grid(self, tables):
        return [self._axis_view(t, a) for a, t in enumerate(tables)]

    def _face_selection(self, axis, high):
        sel = [slice(None)]*3
        sel[axis] = -1 if high else 0
        return tuple(sel)

    def edges(self):
        """Axis edges with the complex edge factor (conservative_helmholtz.py:105-117)."""
        for axis in range(3):
            lo = [slice(None)]*3
            hi = [slice(None)]*3
            lo[axis] = slice(None, -1)
            hi[axis] = slice(1, None)
            lo, hi = tuple(lo), tuple(hi)
            b, c_ = [a for a in range(3) if a != axis]
            s_axis = self._axis_view(self.s1d[axis], axis)
            mid = 0.5*(s_axis[lo] + s_axis[hi])
            factor = (self._axis_view(self.w1d[b], b)*self._axis_view(self.w1d[c_], c_)
                      *self._axis_view(self.s1d[b], b)*self._axis_view(self.s1d[c_], c_)
                      / mid/self.dx**2)
            yield lo, hi, factor

    def diagonal_and_edges(self):
        """Diagonal (with face terms, fixed rows = 1) and the per-edge coefficients."""
        diag = self.mass_weight*self.omega**2/(self.rho*self.c*self.c)
        diag = np.array(np.broadcast_to(diag, self.shape), dtype=np.complex128)
        coefs = []
        for lo, hi, factor in self.edges():
            coef = factor*2.0/(self.rho[lo] + self.rho[hi])
            diag[lo] -= coef
            diag[hi] -= coef
            coefs.append((lo, hi, coef))
        for sel, _, term in self.face_terms:
            diag[sel] += term
        diag[self.fixed] = 1.0
        return diag, coefs

    def matrix(self):
        """complex128 CSR of A (conservative_helmholtz.py:119-149)."""
        n = self.rho.size
        ids = np.arange(n).reshape(self.shape, order="F")
        diag, coefs = self.diagonal_and_edges()
        rows, cols, vals = [], [], []
        for lo, hi, coef in coefs:
            for a, b in ((lo, hi), (hi, lo)):
                keep = ~self.fixed[a]
                rows.append(ids[a][keep])
                cols.append(ids[b][keep])
                vals.append(np.broadcast_to(coef, ids[a].shape)[keep])
        rows.append(ids.ravel())
        cols.append(ids.ravel())
        vals.append(diag.ravel())
        m = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(n, n)).tocsr()
        m.eliminate_zeros()
        return m

    def material_vjp(self, field, adjoint):
        """dL = -Re(lambda^H dA u) (conservative_helmholtz.py:151-175)."""
        u = np.asarray(field).reshape(self.shape, order="F")
        lam = np.array(np.asarray(adjoint).reshape(self.shape, order="F"))
        lam[self.fixed] = 0
        mass = self.mass_weight*self.omega**2/(self.rho*self.c**2)
        prod = np.conj(lam)*u*mass
        g_rho = np.real(prod/self.rho)
        g_c = np.real(2*prod/self.c)
        for lo, hi, factor in self.edges():
            d_coef = -2*factor/(self.rho[lo] + self.rho[hi])**2
            contrib = -np.real((np.conj(lam[lo]) - np.conj(lam[hi]))*(u[hi] - u[lo])*d_coef)
            g_rho[lo] += contrib
            g_rho[hi] += contrib
        for sel, kind, term in self.face_terms:
            if kind == "open":
                contrib = np.real(np.conj(lam[sel])*u[sel]*term)
                g_rho[sel] += contrib/self.rho[sel]
                g_c[sel] += contrib/self.c[sel]
        return g_rho, g_c

    def pml_mask(self):
        sx, sy, sz = self._grid(self.s1d)
        return np.broadcast_to((np.imag(sx) + np.imag(sy) + np.imag(sz)) > 0, self.shape)
