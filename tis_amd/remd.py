"""Temperature replica exchange over replicas held in one device batch.

Replaces the replica loop of energy_allrep.F90:120-230 (each replica's
energy at its own and at its exchange partner's temperature),
simu_replica_exchange.F90:120-302 (Metropolis test of Sugita et al. 2000
eq. 14 between alternating odd/even label pairs, label permutation,
velocity rescale sqrt(T_new/T_old)) and the parameter refresh of
time_integral_post.F90:264-370 (new T-dependent coefficients and Langevin
constants).  Coordinates never leave the device: a replica keeps its slot
and changes its temperature label.
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import rng
from .sysdesc import E_TOTAL
from .topology import BOLTZ_KCAL_MOL, set_temperature


class ReplicaLadder:
    """slots[k]: engine system index of the replica that starts at label k;
    temps ascending.  base: the System all replicas share (topology)."""

    def __init__(self, eng, slots, base, temps, ionic_strength=0.150, seed=1, key=0):
        if any(temps[k] >= temps[k + 1] for k in range(len(temps) - 1)):
            raise ValueError("temperatures must ascend")
        if len(slots) != len(temps):
            raise ValueError("one slot per temperature")
        self.eng = eng
        self.slots = list(slots)
        self.temps = [float(t) for t in temps]
        self.ionic = float(ionic_strength)
        self.seed, self.key = int(seed), int(key)
        self.label_sys = []
        for t in self.temps:
            s = copy.deepcopy(base)
            set_temperature(s, t, self.ionic)
            self.label_sys.append(s)
        self.rep2lab = list(range(len(temps)))   # replica (slot position) -> label
        self.lab2rep = list(range(len(temps)))
        self.parity = 0                          # 0: pairs (1,2),(3,4)..; 1: (2,3),(4,5).. (1-based labels)
        self.attempts = 0
        self.accepted = 0
        self.tried = 0
        # every temperature's coefficients live on the device as a label
        for lab, s in enumerate(self.label_sys):
            eng.label_define(lab, self.slots[0], s)
        eng.label_apply(self.slots, self.rep2lab)

    def partner(self, lab):
        """get_pair of var_replica.F90:158-162 for a 1-D ladder."""
        j = lab + 1 if (lab % 2) == self.parity else lab - 1
        return j if 0 <= j < len(self.temps) else -1

    def energies(self):
        """(E at own T, E at partner T) per replica (energy_allrep.F90:123-216)."""
        own = self.eng.energy_systems(self.slots)[:, E_TOTAL].copy()
        part = np.zeros_like(own)
        moved = [r for r, lab in enumerate(self.rep2lab) if self.partner(lab) >= 0]
        if moved:
            self.eng.label_apply([self.slots[r] for r in moved], [self.partner(self.rep2lab[r]) for r in moved])
            e = self.eng.energy_systems([self.slots[r] for r in moved])
            for k, r in enumerate(moved):
                part[r] = e[k, E_TOTAL]
            self.eng.label_apply([self.slots[r] for r in moved], [self.rep2lab[r] for r in moved])
        return own, part

    def attempt(self):
        """One exchange attempt over every pair of the current parity.
        Returns the list of (label_i, label_j, delta, accepted)."""
        own, part = self.energies()
        out = []
        nrand = 0
        scale = {}
        for li in range(len(self.temps)):
            lj = self.partner(li)
            if lj < li:
                continue
            ri, rj = self.lab2rep[li], self.lab2rep[lj]
            tm, tn = self.temps[li], self.temps[lj]
            delta = ((part[rj] - own[ri]) / tm - (own[rj] - part[ri]) / tn) / BOLTZ_KCAL_MOL
            acc = delta < 0.0 or rng.uniform(self.seed, self.key, self.attempts, 0, nrand) <= math.exp(-delta)
            nrand += 1
            if acc:
                self.rep2lab[ri], self.rep2lab[rj] = lj, li
                self.lab2rep[li], self.lab2rep[lj] = rj, ri
                f = math.sqrt(tn / tm)
                scale[ri], scale[rj] = f, 1.0 / f
                self.accepted += 1
            self.tried += 1
            out.append((li, lj, float(delta), bool(acc)))
        if scale:
            rs = sorted(scale)
            self.eng.label_apply([self.slots[r] for r in rs], [self.rep2lab[r] for r in rs], [scale[r] for r in rs])
        self.parity ^= 1
        self.attempts += 1
        return out

    @property
    def acceptance(self):
        return self.accepted / self.tried if self.tried else 0.0
