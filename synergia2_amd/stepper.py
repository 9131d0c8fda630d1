"""Split-operator stepping of a FODO channel around the space-charge kick.

Restates, for a lattice of drifts and quadrupoles:
  Split_operator_stepper::apply_impl  (simulation/split_operator_stepper.cc:132-261):
      num_steps equal steps; each = first half-step of sliced elements,
      the collective kick, second half-step;
  Step::apply                         (simulation/step.cc:44-78):
      kick time = step_length / (beta c), time_fraction 1.0.
The maps of a second half-step and the next first half-step are applied as
one fused device program (same slices, same order, each particle read and
written once).
"""
import math

import numpy as np

import synergia2_amd as sc


def fodo_channel(ncells=4, ld=0.5, lq=0.2, k1=0.6):
    """Cells of (drift ld, focusing quad lq k1, drift ld, defocusing quad lq -k1)."""
    cell = [(sc.MAP_DRIFT, ld, 0.0), (sc.MAP_QUADRUPOLE, lq, +k1),
            (sc.MAP_DRIFT, ld, 0.0), (sc.MAP_QUADRUPOLE, lq, -k1)]
    return cell * ncells


def _slices(elements, s0, s1, tol=1e-12):
    """Element slices covering [s0, s1) of the lattice (kind, length, k1)."""
    out = []
    pos = 0.0
    for kind, length, k1 in elements:
        a, b = max(s0, pos), min(s1, pos + length)
        if b - a > tol:
            out.append((kind, b - a, k1))
        pos += length
    return out


def split_operator_steps(elements, num_steps):
    """[(first_half_slices, second_half_slices, step_length)] for num_steps
    equal steps over the lattice."""
    total = sum(e[1] for e in elements)
    step = total / num_steps
    half = 0.5 * step
    steps = []
    for i in range(num_steps):
        left = i * step
        steps.append((_slices(elements, left, left + half),
                      _slices(elements, left + half, left + step), step))
    return steps


def split_operator_steps_elements(elements, steps_per_element=1):
    """Split_operator_stepper_elements::apply_impl
    (simulation/split_operator_stepper_elements.cc:4-51): every element of
    non-zero length gets steps_per_element equal steps, each = first half of
    the step's slice of that element, the kick, second half.  Same triple
    format as split_operator_steps."""
    if steps_per_element < 1:
        raise ValueError("Split_operator_stepper_elements: steps_per_element must be >= 1")
    steps = []
    for kind, length, k1 in elements:
        if length == 0.0:
            continue
        step = length / steps_per_element
        for i in range(steps_per_element):
            steps.append(([(kind, 0.5 * step, k1)], [(kind, 0.5 * step, k1)], step))
    return steps


def halo_mix_parts(n, chol, seed, halo_fraction=0.1, halo_radius=3.0):
    """(7, n) host particles: a 6D Gaussian (mean 0, packed lower factor
    chol) of which the last halo_fraction of the slots are replaced by a
    transverse tail population, uniform inside the 4D (x, x', y, y')
    hyper-ellipsoid of radius halo_radius in the Gaussian's normalised
    coordinates (a K-V-like uniform ellipse at halo_radius sigma), with
    Gaussian (cdt, dp/p).  numpy Generator seeded with `seed`; id = slot."""
    rng = np.random.default_rng(seed)
    L = np.zeros((6, 6))
    q = 0
    for r in range(6):
        for c in range(r + 1):
            L[r, c] = chol[q]
            q += 1
    u = rng.standard_normal((6, n))
    nh = int(round(halo_fraction * n))
    if nh:
        g = rng.standard_normal((4, nh))
        g /= np.linalg.norm(g, axis=0)
        rad = halo_radius * rng.uniform(0.0, 1.0, nh) ** 0.25  # uniform in the 4-ball
        u[:4, n - nh:] = g * rad
    p = np.zeros((7, n))
    p[:6] = L @ u
    p[6] = np.arange(n)
    return p


def cell_matrix(elements, plane):
    """Linear thick-lens transfer matrix of an element list (plane 0 = x,
    1 = y; a focusing k1 > 0 focuses x)."""
    M = np.eye(2)
    for kind, length, k1 in elements:
        if kind == sc.MAP_DRIFT or k1 == 0.0:
            R = np.array([[1.0, length], [0.0, 1.0]])
        else:
            k = k1 if plane == 0 else -k1
            if k > 0:
                w = math.sqrt(k)
                R = np.array([[math.cos(w * length), math.sin(w * length) / w],
                              [-w * math.sin(w * length), math.cos(w * length)]])
            else:
                w = math.sqrt(-k)
                R = np.array([[math.cosh(w * length), math.sinh(w * length) / w],
                              [w * math.sinh(w * length), math.cosh(w * length)]])
        M = R @ M
    return M


def periodic_twiss(M):
    """(beta, alpha, phase advance) of a stable one-period matrix."""
    cosmu = 0.5 * (M[0, 0] + M[1, 1])
    if abs(cosmu) >= 1:
        raise ValueError("unstable cell")
    sinmu = math.copysign(math.sqrt(1 - cosmu * cosmu), M[0, 1])
    beta = M[0, 1] / sinmu
    alpha = (M[0, 0] - M[1, 1]) / (2 * sinmu)
    return beta, alpha, math.acos(cosmu)


def matched_chol(elements_one_cell, emit, sigma_cdt, sigma_dpop):
    """Packed 6x6 lower-triangular factor of a Gaussian matched to the cell
    (uncoupled planes, rms geometric emittance emit, or (emit_x, emit_y))."""
    L = np.zeros(21)
    emits = tuple(emit) if np.ndim(emit) else (emit, emit)
    for plane in (0, 1):
        beta, alpha, _ = periodic_twiss(cell_matrix(elements_one_cell, plane))
        emit = emits[plane]
        r0 = 2 * plane
        a = math.sqrt(emit * beta)
        L[r0 * (r0 + 1) // 2 + r0] = a
        r1 = r0 + 1
        L[r1 * (r1 + 1) // 2 + r0] = -alpha * math.sqrt(emit / beta)
        L[r1 * (r1 + 1) // 2 + r1] = math.sqrt(emit / beta)
    L[4 * 5 // 2 + 4] = sigma_cdt
    L[5 * 6 // 2 + 5] = sigma_dpop
    return L


def matched_chol_sigma(elements_one_cell, sigma_x, sigma_y, sigma_cdt, sigma_dpop):
    """Gaussian matched to the cell with the given rms beam sizes at the cell
    entrance: emit_plane = sigma_plane^2 / beta_plane."""
    bx = periodic_twiss(cell_matrix(elements_one_cell, 0))[0]
    by = periodic_twiss(cell_matrix(elements_one_cell, 1))[0]
    return matched_chol(elements_one_cell, (sigma_x ** 2 / bx, sigma_y ** 2 / by), sigma_cdt,
                        sigma_dpop)


class Propagator:
    """Tracks one bunch through the split-operator steps of a lattice (one
    turn = num_steps steps), with the space-charge operator as the
    collective kick (Propagator::propagate / do_step,
    simulation/propagator.cc:36-118,157-283, without diagnostics)."""

    def __init__(self, elements, num_steps, bunch, operator, stream=None, steps=None):
        self.bunch = bunch
        self.op = operator
        self.stream = stream
        self.steps = steps if steps is not None else split_operator_steps(elements, num_steps)
        self.kick_dt = [st[2] / (bunch.beta * sc.C_LIGHT) for st in self.steps]
        p, m = bunch.p, bunch.mass
        self.first = [sc.make_slices(st[0], p, m) for st in self.steps]
        self.last = [sc.make_slices(st[1], p, m) for st in self.steps]
        # second half of step i fused with the first half of step i+1
        self.between = []
        for i, st in enumerate(self.steps):
            nxt = self.steps[(i + 1) % len(self.steps)][0]
            prog = list(st[1]) + list(nxt)
            assert len(prog) <= sc.MAX_SLICES
            self.between.append(sc.make_slices(prog, p, m))
        self.i = 0

    def run(self, nsteps, sync=False, fuse=True):
        """nsteps complete steps: half maps, kick, half maps each.  With
        fuse, the maps after each kick run inside the kick kernel."""
        ns = len(self.steps)
        for j in range(nsteps):
            k = self.i % ns
            if j == 0:
                sc.apply_maps(self.bunch, self.first[k], self.stream)
            after = self.between[k] if j < nsteps - 1 else self.last[k]
            if fuse:
                self.op.set_post_kick_maps(after)
                self.op.apply_bunch(self.bunch, self.kick_dt[k], stream=self.stream, sync=False)
            else:
                self.op.set_post_kick_maps(None)
                self.op.apply_bunch(self.bunch, self.kick_dt[k], stream=self.stream, sync=False)
                sc.apply_maps(self.bunch, after, self.stream)
            self.i += 1
        self.op.set_post_kick_maps(None)
        if sync:
            self.op.sync()
