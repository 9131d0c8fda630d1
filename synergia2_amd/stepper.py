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

    def __init__(self, elements, num_steps, bunch, operator, stream=None):
        self.bunch = bunch
        self.op = operator
        self.stream = stream
        self.steps = split_operator_steps(elements, num_steps)
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
