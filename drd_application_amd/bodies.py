"""Rigid particles and elastic fibers on the host, coupled to the device solve.

The reference moves its bodies on the host between the solves of a step
(TP/mcl6.f90:404-415): it spreads each body's force into particle_force1/2
(force_caculation_vel, TP/particles_move.f90:758-839), solves, and moves each
body with its weighted average of the new velocity (update_position_vel,
:842-920).  The grid parts of both are device reductions here
(drd_particle_force, drd_particle_velocity); this module keeps what stays on
the host: the body list, the fibers' internal forces and the position update.

A fiber is a chain of beads whose internal force is the stretching and
bending model of MF/particles_fibers_move.f90:47-107: springs ks (|r_{i+1} -
r_i| - l0) along the chain and a bending term kb / l0 acting on the turning
angle G_i = theta_{i+1} - theta_i - theta0 between consecutive segments.
Particles carry no internal force.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def fiber_forces(pts: np.ndarray, ks: float, kb: float, l0: float, theta0: float) -> np.ndarray:
    """Stretching + bending force on each bead of one chain (n, 2), as
    force_caculation MF/particles_fibers_move.f90:47-107 forms it."""
    n = len(pts)
    F = np.zeros((n, 2))
    if n < 2:
        return F
    d = np.diff(pts, axis=0)                       # segment i: bead i -> i+1
    rd = np.hypot(d[:, 0], d[:, 1])
    t = d / rd[:, None]
    nv = np.stack([-d[:, 1], d[:, 0]], axis=1) / rd[:, None]
    th = np.arctan2(t[:, 1], t[:, 0])
    G = np.zeros(max(n - 2, 0))
    if n >= 3:
        al = th[1:] - th[:-1]
        al = np.where(np.abs(al) >= math.pi / 2, al + 2 * math.pi, al)
        G = al - theta0
    Gp = np.concatenate([G, [0.0, 0.0]])           # G(i) for i <= n-1 (0 beyond the chain)
    fs = ks * (rd - l0)[:, None] * t               # spring of segment i
    F[:-1] += fs
    F[1:] -= fs
    for i in range(n):                             # bending, per bead
        if i == 0:
            fb = -kb / l0 * Gp[0] / rd[0] * nv[0]
        elif i == n - 1:
            fb = -kb / l0 * (Gp[i - 2] / rd[i - 1] * nv[i - 1]) if i >= 2 else np.zeros(2)
        elif i == 1:
            fb = kb / l0 * (Gp[0] / rd[0] * nv[0] - (Gp[1] - Gp[0]) / rd[1] * nv[1])
        elif i == n - 2:
            fb = kb / l0 * ((Gp[i - 1] - Gp[i - 2]) / rd[i - 1] * nv[i - 1] + Gp[i - 1] / rd[i] * nv[i])
        else:
            fb = kb / l0 * ((Gp[i - 1] - Gp[i - 2]) / rd[i - 1] * nv[i - 1] - (Gp[i] - Gp[i - 1]) / rd[i] * nv[i])
        F[i] += fb
    return F


@dataclass
class Bodies:
    """Particles (x, y, radius) and fibers (bead chains of one bead radius).
    Fiber beads keep unwrapped x (a chain may straddle x = lx); the table
    wraps them into [0, lx)."""

    particles: np.ndarray                       # (np, 3)
    lx: float
    fibers: list = field(default_factory=list)  # each (nb, 2)
    bead_radius: float = 0.0
    ks: float = 1.0e3
    kb: float = 1.0
    l0: float = 0.0
    theta0: float = 0.0

    @property
    def count(self) -> int:
        return len(self.particles) + sum(len(f) for f in self.fibers)

    def table(self) -> np.ndarray:
        """(n, 5) rows (x, y, r, fx, fy): particles, then the fibers' beads in
        chain order, with the fibers' internal forces."""
        rows = [np.concatenate([self.particles, np.zeros((len(self.particles), 2))], axis=1)]
        for f in self.fibers:
            F = fiber_forces(f, self.ks, self.kb, self.l0, self.theta0)
            xw = np.mod(f[:, :1], self.lx)
            rows.append(np.concatenate([xw, f[:, 1:], np.full((len(f), 1), self.bead_radius), F], axis=1))
        return np.concatenate(rows, axis=0)

    def spread(self, solver, lx: float, ly: float) -> np.ndarray:
        """particle_force1/2 on the device from the current bodies; returns the table used."""
        t = self.table()
        solver.particle_force(t, lx, ly)
        return t

    def advance(self, solver, dt: float, lx: float, ly: float) -> tuple:
        """update_position_vel: every body moves with its weighted average of the
        solver's velocity; x wraps periodically.  Returns (u_aver, v_aver)."""
        t = self.table()
        ua, va = solver.particle_velocity(t, lx, ly)
        k = len(self.particles)
        self.particles[:, 0] = np.mod(self.particles[:, 0] + dt * ua[:k], lx)
        self.particles[:, 1] += dt * va[:k]
        for f in self.fibers:
            n = len(f)
            f[:, 0] += dt * ua[k:k + n]
            f[:, 1] += dt * va[k:k + n]
            k += n
        return ua, va
