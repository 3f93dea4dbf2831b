"""The one-time Galerkin rescale of the pressure solve (rbgs_rescale_after; no
reference counterpart).  The oracle's factor is checked against the closed
form for a scaled exact solution and against a factor formed here from the
system rebuilt out of the oracle's rows; the rescaled solve is checked against
a solve stopped before the rescale, rescaled here in float32 exactly as the
pressure store does it, and continued from that field."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from scenes import SceneArrays, block_particles, box_settings

FP = 2.0 ** 18


def _pool_settings(**kw):
    s = box_settings(n=16, cup=True, tolerance=1e-5, **kw)
    return s


def _pool(orc, s):
    n = s.grid_dims[0]
    r_cells = 0.42 * n - 1.5

    def inside(ix, iy, iz):
        return (ix + 0.5 - n / 2) ** 2 + (iz + 0.5 - n / 2) ** 2 <= r_cells ** 2

    parts = block_particles(s, (0, 3, 0), (n, 9, n), inside=inside)
    sc = SceneArrays(parts)
    return orc.OracleSim(s, sc.scene), sc


def _system(s, o):
    """A and b of the oracle's pressure rows (fluid cells with a non-empty stencil)."""
    w, count, c = o.read_system()
    _, kind = o.read_pressure()
    fluid = np.nonzero((kind >= 1) & (kind <= 3) & (count > 0))[0]
    pos = {int(i): k for k, i in enumerate(fluid)}
    n = s.grid_dims[0]
    off = [-1, 1, -n, n, -n * n, n * n]
    rows, cols, vals = [], [], []
    for k, i in enumerate(fluid):
        rows.append(k); cols.append(k); vals.append(float(count[i]))
        for d in range(6):
            j = int(i) + off[d]
            if w[i, d] > 0 and j in pos:
                rows.append(k); cols.append(pos[j]); vals.append(-float(w[i, d]))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(fluid.size, fluid.size))
    return A, -c[fluid].astype(np.float64), fluid


def _factor(A, b, p):
    return float(p @ b) / float(p @ (A @ p))


def _round_factor(a):
    return np.float32(np.rint(a * 2.0 ** 20) / 2.0 ** 20)


@pytest.mark.parametrize("scale", [0.8, 1.25])
def test_factor_of_a_scaled_solution_is_its_inverse(orc, scale):
    s = _pool_settings()
    o, _sc = _pool(orc, s)
    o.debug_step_until(1e-3, "rbgs")
    A, b, fluid = _system(s, o)
    p_star = spla.spsolve(A.tocsc(), b)
    p_fixed = np.zeros(o.nc, np.int32)
    # quantised as the pressure store writes it, so every value is a float32 integer
    p_fixed[fluid] = np.trunc((scale * p_star).astype(np.float32) * np.float32(FP)).astype(np.int32)
    o.debug_set_pressure(p_fixed)
    a_orc = o.rescale_factor()
    a_np = _factor(A, b, p_fixed[fluid] / FP)
    assert abs(a_orc - a_np) <= 1e-12 * abs(a_np)
    # closed form: alpha (s p*) = 1 / s, up to the 2^-18 quantisation of the field
    assert abs(a_np - 1.0 / scale) <= 1e-5
    o.close()


def _run(orc, s, warm=None):
    o, _sc = _pool(orc, s)
    if warm is not None:
        o.set_warm_start(warm)
    o.debug_step_until(1e-3, "rbgs")
    return o


def test_rescaled_solve_equals_a_solve_rescaled_by_hand(orc):
    m, every = 16, 8
    # A: m untested iterations, the rescale, then `every` more
    sa = _pool_settings()
    sa.rbgs_tolerance = 0.0
    sa.rbgs_rescale_after = m
    sa.rbgs_max_iters = m + every
    oa = _run(orc, sa)
    assert oa.solve_stats().iterations == m + every
    pa, _ = oa.read_pressure()
    # B: the same m iterations and nothing else
    sb = _pool_settings()
    sb.rbgs_tolerance = 0.0
    sb.rbgs_max_iters = m
    ob = _run(orc, sb)
    pb, _ = ob.read_pressure()
    A, b, fluid = _system(sb, ob)
    a_np = _factor(A, b, pb[fluid] / FP)
    assert abs(ob.rescale_factor() - a_np) <= 1e-12 * abs(a_np)
    assert 0.0625 < a_np < 16.0 and abs(a_np - 1.0) > 1e-3, "the rescale should change this field"
    alpha = _round_factor(a_np)
    # the rescale in float32 as the pressure store does it: clamp, scale to fixed point, truncate
    pf = pb.astype(np.float32) * np.float32(2.0 ** -18)
    v = alpha * pf
    lim = np.float32(8191.0)
    v = np.minimum(np.maximum(v, -lim), lim)
    p_scaled = np.trunc(v * np.float32(FP)).astype(np.int32)
    _, kind = ob.read_pressure()
    is_fluid = (kind >= 1) & (kind <= 3)
    warm = np.where(is_fluid, p_scaled.astype(np.float32) * np.float32(2.0 ** -18), np.float32(0.0))
    # C: continue from the hand-rescaled field for `every` iterations
    sc_ = _pool_settings()
    sc_.rbgs_tolerance = 0.0
    sc_.rbgs_max_iters = every
    sc_.warm_start = 2  # CS_WARM_EXTERNAL
    sc_.warm_fallback_residual = 1e30
    oc = _run(orc, sc_, warm=warm.astype(np.float32))
    assert oc.solve_stats().iterations == every
    pc, _ = oc.read_pressure()
    np.testing.assert_array_equal(pa, pc)
    for o in (oa, ob, oc):
        o.close()


def test_rescale_keeps_the_tolerance_and_counts_its_iterations(orc):
    s = _pool_settings()
    s.rbgs_tolerance = 1e-5
    s.rbgs_max_iters = 4000
    plain = _run(orc, s)
    s2 = _pool_settings()
    s2.rbgs_tolerance = 1e-5
    s2.rbgs_max_iters = 4000
    s2.rbgs_rescale_after = 16
    resc = _run(orc, s2)
    a, b = plain.solve_stats(), resc.solve_stats()
    assert a.converged == 1 and b.converged == 1
    assert b.final_residual <= 1e-5
    assert b.iterations >= 16 + 8 and b.iterations % 8 == 0
    assert b.initial_residual == a.initial_residual
    plain.close()
    resc.close()


def test_invalid_rescale_settings_are_rejected(lib):
    """The library's settings check (run by every entry point that takes
    settings; cs_generate_sdf runs on the host) refuses a rescale that is not a
    multiple of 4 and of the check interval, leaves no tested chunk under the
    cap, or asks for it in fixed-pairs mode."""
    import coffee_sim_amd as cs

    ok = _pool_settings()
    ok.rbgs_rescale_after = 16
    ok.rbgs_max_iters = 100
    cs.generate_sdf(ok)
    for after, every, cap, mode in ((6, 2, 100, 1), (12, 8, 100, 1), (96, 8, 100, 1), (16, 8, 100, 0)):
        s = _pool_settings()
        s.rbgs_rescale_after = after
        s.rbgs_check_every = every
        s.rbgs_max_iters = cap
        s.rbgs_mode = mode
        with pytest.raises(cs.CoffeeSimError):
            cs.generate_sdf(s)
