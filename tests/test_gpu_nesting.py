"""Nest hierarchy driver (vibe_mi.nesting.NestedRun) on the GPU.

* A two-level hierarchy (parent, one ratio-3 nest, one nest inside it)
  advanced by the batched driver matches the same schedule written out by
  hand with single-domain steps and the nest entry points, bit for bit.
* The library's graph-captured parent step (vibe_mi_batch_nested_step)
  equals the same schedule issued sub-step by sub-step from Python, bit for
  bit, over four parent steps (every buffer-rotation state of the parent),
  one-way and two-way, FP64 and FP32-mixed.
* A uniform flow over flat terrain is an exact steady state of every
  domain, so parent and nests stay uniform through the relaxation (to the
  rounding of the relaxation's weighted sum).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def _hier(sigma=300.0, theta_amp=0.5, sponge_alpha=0.1):
    from vibe_mi import configs
    return configs.nest_hierarchy(pnx=48, pny=42, nz=8, dx=3000.0, sigma=sigma, theta_amp=theta_amp,
                                  nests=[(27, 24, 0, (6, 6)), (18, 18, 1, (2, 3))], steps=2,
                                  sponge_alpha=sponge_alpha)


def test_nested_run_matches_hand_schedule(vmi):
    from vibe_mi.nesting import NestedRun
    cases, links = _hier()
    run = NestedRun(cases, links, precision=vmi.FP64)
    run.step(2)
    got = [m.get_state() for m in run.members]
    run.close()

    P, A, B = (vmi.Domain(c, vmi.FP64) for c in cases)
    dA = vmi.NestDesc(3, links[0][3], links[0][4], 5, 1.0)
    dB = vmi.NestDesc(3, links[1][3], links[1][4], 5, 1.0)
    vmi.nest_interpolate(A, P, dA)
    vmi.nest_interpolate(B, A, dB)
    dt0, ns = cases[0].dt, cases[0].nsub
    for _ in range(2):            # parent steps
        for s in range(9):        # finest sub-steps
            if s == 0:
                P.step(dt0, ns, 1)
            if s % 3 == 0:
                A.step(dt0 / 3, ns, 1)
            B.step(dt0 / 9, ns, 1)
            if s % 3 == 0:
                vmi.nest_relax(A, P, dA, (s // 3 + 1) / 3)
            vmi.nest_relax(B, A, dB, (s % 3 + 1) / 3)
    want = [d.get_state() for d in (P, A, B)]
    for d in (P, A, B):
        d.close()
    for m, (g, w) in enumerate(zip(got, want)):
        for f in FIELDS:
            assert np.array_equal(g[f], w[f]), (m, f, float(np.max(np.abs(g[f] - w[f]))))


def test_uniform_flow_stays_exact_through_nesting(vmi):
    from vibe_mi.nesting import NestedRun
    cases, links = _hier(sigma=0.0, theta_amp=0.0, sponge_alpha=0.0)  # the sponge damps u toward 0
    # FP64: the uncontracted arithmetic keeps the uniform state exact (the
    # FP32 batched kernels use fused multiply-adds, which round the two sides
    # of a cancelling flux difference differently)
    run = NestedRun(cases, links, precision=vmi.FP64)
    assert run.cell_updates_per_parent_step() == sum(c.ncell * 3 ** lv for c, lv in zip(cases, (0, 1, 2)))
    run.step(1)
    for m, c in zip(run.members, cases):
        st = m.get_state()
        # the relaxation (1 - a) u + a p may round u by an ulp
        assert np.max(np.abs(st["u"] - 10.0)) < 1e-12, c.name
        for f in ("v", "w", "rho", "theta", "pi"):
            assert np.all(st[f] == 0.0), (c.name, f)
    run.close()


def test_batch_nest_relax_rejects_bad_members(vmi):
    cases, links = _hier()
    b = vmi.Batch(cases, vmi.FP64)
    d = vmi.NestDesc(3, links[0][3], links[0][4], 5, 1.0)
    for c, p in ((1, 1), (-1, 0), (1, 7)):
        with pytest.raises(vmi.VibeError) as e:
            vmi.batch_nest_relax(b, c, p, d, 0.5)
        assert e.value.code == vmi.VIBE_C_ERR_BAD_ARG
    b.close()


@pytest.mark.parametrize("prec_name", ["FP64", "MIXED", "MIXED_COMP"])
@pytest.mark.parametrize("two_way", [False, True])
def test_graph_parent_step_equals_substep_schedule(vmi, prec_name, two_way):
    from vibe_mi.nesting import NestedRun
    prec = getattr(vmi, prec_name)
    cases, links = _hier()
    runs = [NestedRun(cases, links, precision=prec, two_way=two_way, native=nat) for nat in (True, False)]
    for r in runs:
        r.step(1)
        r.step(3)
    for m in range(len(cases)):
        a, b = runs[0].members[m].get_state(), runs[1].members[m].get_state()
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), (prec_name, two_way, m, f)
    for r in runs:
        r.close()
