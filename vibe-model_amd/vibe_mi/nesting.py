"""Nest hierarchy driver: a parent and its nests advanced together, one
batched launch sequence per finest-level sub-step, with ratio-r time
refinement and Davies lateral boundaries from each nest's parent.

Per parent step (levels 0..L, ratio r): the finest level takes r**L
sub-steps; at sub-step s the domains of level l are active when
s % r**(L - l) == 0, so level l advances r**l times with dt / r**l.  All
active domains run in the same batched launches (vibe_mi.Batch).  After each
sub-step every nest that just stepped is relaxed toward its parent,
interpolated bilinearly in space and linearly in time across the parent's
current step (frac = (m + 1) / r for the nest's m-th step inside it), coarse
levels first, enqueued on the batch stream with no host round trip.  With
two_way, after the parent step's last sub-step (all levels at the same time)
every nest feeds its mean state back onto the parent cells it covers, finest
level first (batch_nest_restrict).  These are the reference's nest operators
(src/grid/nest.cpp:52-128; its driver never calls them, driver.cpp:212-219).

By default the schedule runs in the library (vibe_mi_batch_nested_step): a
whole parent step is one hipGraph launch and a call synchronises once.
native=False issues the same schedule from here, sub-step by sub-step.
"""
from __future__ import annotations

from . import (MIXED, Batch, NestDesc, batch_nest_relax, batch_nest_restrict, batch_nested_step,
               nest_interpolate)


class NestedRun:
    def __init__(self, cases, links, precision: int = MIXED, use_graph: bool = True, zone: int = 5,
                 alpha_max: float = 1.0, two_way: bool = False, native: bool = True):
        self.two_way = two_way
        self.native = native
        self.cases = list(cases)
        self.links = [(c, p, NestDesc(r, i0, j0, zone, alpha_max)) for (c, p, r, i0, j0) in links]
        ratios = {r for (_, _, r, _, _) in links}
        if len(ratios) > 1:
            raise ValueError("one refinement ratio per hierarchy")
        self.r = ratios.pop() if ratios else 1
        self.level = [0] * len(self.cases)
        for c, p, _ in sorted(self.links, key=lambda x: x[0]):
            self.level[c] = self.level[p] + 1
        self.L = max(self.level)
        self.dt0 = self.cases[0].dt
        self.nsub = self.cases[0].nsub
        self.batch = Batch(self.cases, precision, use_graph=use_graph)
        self.order = sorted(self.links, key=lambda x: self.level[x[0]])
        for c, p, d in self.order:  # nests start as their parent's interpolation
            nest_interpolate(self.batch.members[c], self.batch.members[p], d)

    @property
    def members(self):
        return self.batch.members

    def cell_updates_per_parent_step(self) -> int:
        return sum(c.ncell * self.r ** lv for c, lv in zip(self.cases, self.level))

    def _substep_plan(self):
        n = self.r ** self.L
        dts = [self.dt0 / self.r ** lv for lv in self.level]
        plan = []
        for s in range(n):
            active = [1 if s % self.r ** (self.L - lv) == 0 else 0 for lv in self.level]
            relax = []
            for c, p, d in self.order:
                if active[c]:
                    m = (s // self.r ** (self.L - self.level[c])) % self.r
                    relax.append((c, p, d, (m + 1) / self.r))
            plan.append((active, relax))
        return dts, plan

    def step(self, nsteps: int = 1, timed: bool = False) -> float:
        """nsteps parent steps; returns the summed batched-launch time (ms)
        when timed (the relaxation launches are outside it)."""
        if self.native:
            return batch_nested_step(self.batch, self.links, self.two_way, self.dt0, self.nsub, nsteps, timed)
        dts, plan = self._substep_plan()
        ms = 0.0
        for _ in range(nsteps):
            for active, relax in plan:
                if timed:
                    ms += self.batch.step_timed(dts, self.nsub, 1, active)
                else:
                    self.batch.step(dts, self.nsub, 1, active)
                for c, p, d, frac in relax:
                    batch_nest_relax(self.batch, c, p, d, frac)
            if self.two_way:
                for c, p, d in reversed(self.order):
                    batch_nest_restrict(self.batch, c, p, d)
        return ms

    def close(self):
        self.batch.close()
