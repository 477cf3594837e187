"""FP64 restatement of the phase-optimisation losses, basis field and VJP.

Follows ``src/algorithm/training/phase_adjoint_optimizer.py``
(Gaussian target fields :78-139, nearest grid index :142-144,
``AmplitudeFieldLoss.evaluate`` :163-177, ``_focal_term`` :210-223,
``FocalPressureLoss.evaluate`` :244-258, ``FocalContrastLoss.evaluate``
:312-338, background mask :384-414), the ratio form of focal contrast that
the 160^3 gradient workload states (no reference code; pinned by its
piecewise-constant closed form, scale invariance and central differences),
``src/algorithm/training/phase_response_basis.py`` (trilinear
``target_response`` :139-175, ``field`` :177-206, ``phase_vjp`` :208-253),
``src/algorithm/baselines/common.py`` (focal gradient :189-195, metrics
:197-216, the evaluate composition :275-343) and the central-difference
check of ``src/algorithm/baselines/adjoint/__init__.py:44-58``.

Grids are (nx, ny, nz) arrays; flattening is Fortran order.
Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np

EPS = 1.0e-12


def nearest_index(point, origin, dx):
    return tuple(np.rint((np.asarray(point, float) - np.asarray(origin, float))/dx).astype(int))


def gaussian_target(shape, box, origin, dx, points, peak, sigma, background=0.0,
                    target_weight=1.0, background_weight=0.05):
    points = np.asarray(points, dtype=float)
    if points.ndim != 2 or points.shape[1] != 3 or len(points) == 0:
        raise ValueError("target_points must have shape (num_targets, 3)")
    if sigma <= 0:
        raise ValueError("target sigma must be positive")
    target = np.full(shape, float(background))
    weight = np.full(shape, float(background_weight))
    axes = [np.linspace(origin[a], origin[a] + box[a], shape[a]) for a in range(3)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    for p in points:
        r2 = (gx - p[0])**2 + (gy - p[1])**2 + (gz - p[2])**2
        g = np.exp(-0.5*r2/(sigma*sigma))
        g = g/g[nearest_index(p, origin, dx)]
        target = np.maximum(target, background + (peak - background)*g)
        weight = np.maximum(weight, background_weight + (target_weight - background_weight)*g)
    return target, weight


def background_mask(shape, box, points, radius, exclusion_layers):
    """Grid from 0 (origin ignored, as in the reference) minus source layers and target balls."""
    if radius <= 0:
        raise ValueError("training.target_radius must be positive")
    axes = [np.linspace(0.0, box[a], shape[a]) for a in range(3)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    mask = np.ones(shape, dtype=bool)
    layers = max(int(exclusion_layers), 0)
    if layers:
        mask[:, :, :layers] = False
    for p in np.asarray(points, dtype=float):
        mask[(gx - p[0])**2 + (gy - p[1])**2 + (gz - p[2])**2 <= radius*radius] = False
    return mask


def field_match(field, target, weight, with_cotangent=True):
    amp = np.abs(field)
    res = amp - target
    wres = weight*res
    loss = 0.5*float(np.vdot(wres, wres).real)
    if not with_cotangent:
        return loss, None
    return loss, (weight*weight*res*field/(amp + EPS)).astype(np.complex128)


def focal_term(field, indices, weight, with_cotangent=True, focal_field=None):
    idx = tuple(np.asarray(indices, dtype=np.intp).T)
    values = field[idx] if focal_field is None else np.asarray(focal_field)
    amp = np.abs(values)
    loss = -weight*np.mean(amp)
    cot = np.zeros(field.shape, dtype=np.complex128) if with_cotangent else None
    if with_cotangent and focal_field is None:
        contrib = -weight/len(indices)*values/(amp + EPS)
        np.add.at(cot, idx, contrib)
    return loss, cot


def focal_pressure(field, indices, target_weight, background_weight, with_cotangent=True,
                   focal_field=None):
    amp = np.abs(field)
    loss, cot = focal_term(field, indices, target_weight, with_cotangent, focal_field)
    if background_weight > 0:
        loss = loss + 0.5*background_weight*float(np.mean(amp*amp))
        if with_cotangent:
            cot = cot + background_weight/field.size*field
    return float(loss), cot


def focal_contrast(field, indices, target_weight, sidelobe_weight, temperature, mask,
                   with_cotangent=True, focal_field=None):
    amp = np.abs(field)
    loss, cot = focal_term(field, indices, target_weight, with_cotangent, focal_field)
    bg = amp[mask]
    top = float(np.max(bg))
    ex = np.exp((bg - top)/temperature)
    prob = ex/np.sum(ex)
    smooth = top + temperature*(np.log(np.sum(ex)) - np.log(bg.size))
    loss = float(loss) + sidelobe_weight*float(smooth)
    if not with_cotangent:
        return loss, None
    cot[mask] += sidelobe_weight*prob*field[mask]/(bg + EPS)
    return loss, cot


def ratio_masks(shape, box, points, radius):
    """(ball, rest) of the focal contrast ratio: nodes within `radius` of a
    target (grid from 0, as ``background_mask``) and the nodes of the upper
    half z >= Lz/2 outside every ball."""
    axes = [np.linspace(0.0, box[a], shape[a]) for a in range(3)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    ball = np.zeros(shape, dtype=bool)
    for p in np.asarray(points, dtype=float).reshape(-1, 3):
        ball |= (gx - p[0])**2 + (gy - p[1])**2 + (gz - p[2])**2 <= radius*radius
    upper = np.zeros(shape, dtype=bool)
    upper[:, :, shape[2]//2:] = True
    return ball, upper & ~ball


def focal_contrast_ratio(field, ball, rest, with_cotangent=True):
    """The focal contrast of the 160^3 gradient workload as a ratio: mean |u|^2
    over the target ball divided by mean |u|^2 over the rest of the upper half.
    No reference code exists for this form (its focal_contrast is the
    smooth-max form above); with a = mean_ball |u|^2, b = mean_rest |u|^2 and
    d|u|^2 = Re((2u)^H du): c = 2u/(n_ball b) on the ball,
    -2 a u/(n_rest b^2) on the rest."""
    p2 = np.abs(field)**2
    na, nb = int(np.count_nonzero(ball)), int(np.count_nonzero(rest))
    a = float(np.sum(p2[ball]))/na
    b = float(np.sum(p2[rest]))/nb
    if not with_cotangent:
        return a/b, None
    cot = np.zeros(field.shape, dtype=np.complex128)
    cot[ball] = 2.0*field[ball]/(na*b)
    cot[rest] = -2.0*a*field[rest]/(nb*b*b)
    return a/b, cot


def trilinear_rows(points, shape, origin, dx):
    """(rows, weights) of the 8-corner trilinear stencil (phase_response_basis.py:165-174)."""
    shape = np.asarray(shape)
    pos = (np.asarray(points, float) - np.asarray(origin, float))/dx
    low = np.minimum(np.floor(pos).astype(int), shape - 2)
    frac = pos - low
    rows, weights = [], []
    for corner in np.ndindex(2, 2, 2):
        corner = np.asarray(corner)
        idx = low + corner
        weights.append(np.prod(np.where(corner, frac, 1.0 - frac), axis=1))
        rows.append(idx[:, 0] + shape[0]*(idx[:, 1] + shape[1]*idx[:, 2]))
    return np.stack(rows, 1), np.stack(weights, 1)


def target_response(basis, points, shape, origin, dx):
    rows, w = trilinear_rows(points, shape, origin, dx)
    out = np.zeros((len(rows), basis.shape[1]), dtype=np.complex128)
    for c in range(8):
        out += w[:, c, None]*basis[rows[:, c]]
    return out


def basis_field(basis, phases, shape):
    return (basis @ np.exp(1j*np.asarray(phases, float))).reshape(shape, order="F")


def phase_vjp(basis, cotangent, phases):
    """Re(conj(G^H c) * i e^{i phi}) (phase_response_basis.py:241-253)."""
    c = np.asarray(cotangent, dtype=np.complex128).reshape(-1, order="F")
    resp = basis.conj().T @ c
    return np.real(np.conj(resp)*(1j*np.exp(1j*np.asarray(phases, float))))


def focal_gradient(response, focal, phases, target_weight):
    """common.py:189-195."""
    cot = -target_weight/len(focal)*focal/(np.abs(focal) + EPS)
    adj = response.conj().T @ cot
    return np.real(adj.conj()*1j*np.exp(1j*np.asarray(phases)))


def central_difference(loss_fn, phases, step, count=8):
    """Central FD on the first min(count, M) emitters with the scaled error of
    adjoint/__init__.py:44-58; returns (fd, index list)."""
    phases = np.asarray(phases, float)
    out = []
    for i in range(min(count, phases.size)):
        e = np.zeros_like(phases)
        e[i] = step
        out.append((loss_fn(phases + e) - loss_fn(phases - e))/(2*step))
    return np.asarray(out)


def scaled_gradient_error(exact, numerical):
    return abs(exact - numerical)/max(abs(exact), abs(numerical), 1.0)


def metrics(field, target, target_amp_values, mask):
    """Subset of ``PhaseProblem.metrics`` (common.py:197-216)."""
    amp = np.abs(field)
    values = np.asarray(target_amp_values)
    mean, peak = float(np.mean(values)), float(np.max(values))
    bg = float(np.max(amp[mask]))
    norm = float(np.linalg.norm(target))
    err = float(np.linalg.norm(amp - target))
    return {"target_mean": mean, "target_min": float(np.min(values)), "background_max": bg,
            "contrast": mean/bg if bg > 0 else None,
            "uniformity": float(np.min(values))/peak if peak > 0 else None,
            "field_relative_l2": err/norm if norm > 0 else None, "field_absolute_l2": err}
