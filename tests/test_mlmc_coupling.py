"""The MLMC coupled stepper (CoupledASVJStepper::multi_step_coupled,
CoupledASVJStepper.hpp:170-517) in the CPU oracle, pinned by replaying its
own draws through the plain step equations.

Every uniform a debug path draws is logged (oracle_debug_*_path), so a test
can rebuild the Gaussians (through the oracle's Acklam inverse normal,
itself pinned bit for bit to the reference's Numerics.hpp) and step the
state with the ASVJ step equations written out here:
  variance (SchemeNV, VolSchemes.hpp:24-66):
      v' = K0 (sqrt(K0 v + K1) + K2 Z)^2 + K1
  log price (ASVJstepper.hpp:134-229, constants init_cir_impl :262-283):
      x' = x + sum_f (C0 + C1 v + C2 v') + sqrt(sum_f C3 (v + v')) G (+ jump)
1. Those equations reproduce the oracle's plain stepper path for path
   (so they are the stepper's own).
2. The coupled stepper's fine path is two plain fine steps on (G1, Z1) and
   (G2, Z2), and its coarse path is ONE plain step of twice the size driven
   by the aggregated increments (:225-349): Zc = (Z1 + Z2)/sqrt2 for NV
   variance, the fine variance copied for NCI / Exact, Gc = (a1 G1 +
   a2 G2)/hypot(a1, a2) with a_i the fine steps' diffusion scales, the coarse
   jump the sum of the fine jumps, the European's one jump shared, and the
   coarse tracker updated on the coarse grid.
3. What makes the telescoping sum unbiased: Zc and Gc are standard normal
   and uncorrelated (conditionally on the Z's, Gc is a unit-norm combination
   of the independent G1, G2), checked on many recorded steps.
"""
import math

import numpy as np
import pytest

from myquantlib_amd import capi
from myquantlib_amd.pricing import Heston, InstrumentSpec, make_model, merton, kou

S0, R, Q = 100.0, 0.03, 0.01
H1 = Heston(2.0, 0.04, 0.3, -0.7, 0.04)
H2 = Heston(1.0, 0.03, 0.25, -0.5, 0.02)
H_HI = Heston(2.0, 0.04, 1.0, -0.7, 0.04)
SQ2 = math.sqrt(2.0)


def consts(hs, schemes, dt):
    """Per factor (C0..C3, K0..K2) for a step dt (init_cir_impl; SchemeNV)."""
    out = []
    rp = (R - Q) / len(hs)
    for h, sch in zip(hs, schemes):
        term = (h.kappa * h.rho / h.sigma - 0.5) * dt * 0.5
        C = ((rp - h.rho * h.kappa * h.theta / h.sigma) * dt, term - h.rho / h.sigma, term + h.rho / h.sigma,
             dt * (1.0 - h.rho * h.rho) * 0.5)
        K = None
        if sch == capi.SCHEME_NV:
            ex = math.exp(-h.kappa * dt * 0.5)
            K = (ex, (h.kappa * h.theta - h.sigma ** 2 * 0.25) * (1.0 - ex) / h.kappa, h.sigma * math.sqrt(dt) * 0.5)
        out.append((C, K))
    return out


def nv(K, v, Z):
    root = math.sqrt(K[0] * v + K[1]) + K[2] * Z
    return K[0] * root * root + K[1]


def x_step(cs, x, v, vn, G):
    """One log-price step; returns (x', diffusion scale a)."""
    d = sum(C[0] + C[1] * a + C[2] * b for (C, _), a, b in zip(cs, v, vn))
    a = math.sqrt(sum(C[3] * (p + q) for (C, _), p, q in zip(cs, v, vn)))
    return x + d + a * G, a


def close(a, b, tol=1e-13):
    return abs(a - b) <= tol * max(1.0, abs(b))


@pytest.mark.parametrize("nf", [1, 2])
@pytest.mark.parametrize("pd", [False, True])
def test_step_equations_reproduce_the_plain_stepper(oracle, nf, pd):
    hs = [H1, H2][:nf]
    sch = [capi.SCHEME_NV] * nf
    m = make_model(nf, hs)
    kind = capi.INSTR_ARITH_ASIAN if pd else capi.INSTR_EUROPEAN
    inst = InstrumentSpec(kind=kind, strikes=[100.0], maturity=1.0)
    steps = 24
    rec, fin, draws = oracle.debug_plain_path(m, sch, inst, steps, 7, S0, R, Q)
    cs = consts(hs, sch, 1.0 / steps)
    x, v = math.log(S0), [h.v0 for h in hs]
    for k in range(steps):
        i0 = int(rec[k, 0])
        G = oracle.normcdfinv(draws[i0])
        vn = [nv(K, vv, oracle.normcdfinv(draws[i0 + 1 + f])) for f, ((_, K), vv) in enumerate(zip(cs, v))]
        x, _ = x_step(cs, x, v, vn, G)
        v = vn
        assert close(x, rec[k, 1]) and close(v[0], rec[k, 2]), (k, x, rec[k, 1], v[0], rec[k, 2])
        if nf > 1:
            assert close(v[1], rec[k, 3])
    assert len(draws) == steps * (1 + nf)


CASES = {
    # name: (model, schemes, instrument kind)
    "heston_nv_european": (make_model(1, [H1]), [capi.SCHEME_NV], capi.INSTR_EUROPEAN),
    "double_heston_nv_asian": (make_model(2, [H1, H2]), [capi.SCHEME_NV, capi.SCHEME_NV], capi.INSTR_ARITH_ASIAN),
    "bates_nci_european": (make_model(1, [H_HI], jump_kind=merton(0.5, -0.1, 0.15)[0],
                                      jump=merton(0.5, -0.1, 0.15)[1]), [capi.SCHEME_NCI], capi.INSTR_EUROPEAN),
    "double_bates_nci_exact_asian": (make_model(2, [H_HI, H2], jump_kind=merton(0.5, -0.1, 0.15)[0],
                                                jump=merton(0.5, -0.1, 0.15)[1]),
                                     [capi.SCHEME_NCI, capi.SCHEME_EXACT], capi.INSTR_ARITH_ASIAN),
    "bates_kou_nv_asian": (make_model(1, [H1], jump_kind=kou(1.0, 0.4, 10.0, 5.0)[0], jump=kou(1.0, 0.4, 10.0, 5.0)[1]),
                           [capi.SCHEME_NV], capi.INSTR_ARITH_ASIAN),
}


# record columns of a coarse step (oracle.debug_coupled_path)
I_G1, I_G2, I_GC, I_A1, I_A2, I_J1, I_J2, I_XF, I_VF, I_XC, I_VC, I_MID = 1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 14


@pytest.mark.parametrize("name", sorted(CASES))
def test_coarse_path_is_one_plain_step_on_aggregated_increments(oracle, name):
    m, sch, kind = CASES[name]
    nf = m.n_factors
    hs = [Heston(m.heston[f].kappa, m.heston[f].theta, m.heston[f].sigma, m.heston[f].rho, m.heston[f].v0)
          for f in range(nf)]
    T, nc = 1.0, 16
    inst = InstrumentSpec(kind=kind, strikes=[100.0], maturity=T)
    pd = kind != capi.INSTR_EUROPEAN
    all_nv = all(s == capi.SCHEME_NV for s in sch)
    jumps = m.jump_kind != capi.JUMP_NONE
    for seed in (3, 11):
        rec, ff, fc, draws = oracle.debug_coupled_path(m, sch, inst, nc, seed, S0, R, Q)
        dtf, dtc = T / (2 * nc), T / nc
        cf, cc = consts(hs, sch, dtf), consts(hs, sch, dtc)
        # the European's single jump is drawn first and shared by both paths
        J = rec[nc, 0] if (jumps and not pd) else 0.0
        xf = xc = math.log(S0) + J
        vf, vc = [h.v0 for h in hs], [h.v0 for h in hs]
        sumf, sumc = 0.5 * S0 * dtf, 0.5 * S0 * dtc   # arithmetic-Asian trapezoids
        for k in range(nc):
            r = rec[k]
            i0 = int(r[0])
            G1, G2 = oracle.normcdfinv(draws[i0]), oracle.normcdfinv(draws[i0 + 1])
            assert r[I_G1] == G1 and r[I_G2] == G2
            mid, nxt = list(r[I_MID:I_MID + nf]), list(r[I_VF:I_VF + nf])
            Z1 = Z2 = None
            if all_nv:
                # fine step 1 draws Z1 per factor right after G1, G2
                Z1 = [oracle.normcdfinv(draws[i0 + 2 + f]) for f in range(nf)]
                assert all(close(a, b) for a, b in zip([nv(K, v, z) for (_, K), v, z in zip(cf, vf, Z1)], mid))
                if not jumps:
                    Z2 = [oracle.normcdfinv(draws[i0 + 2 + nf + f]) for f in range(nf)]
                    assert all(close(a, b) for a, b in zip([nv(K, v, z) for (_, K), v, z in zip(cf, mid, Z2)], nxt))
            # fine path: two plain fine steps (+ their jumps)
            xm, a1 = x_step(cf, xf, vf, mid, G1)
            xm += r[I_J1]
            xf, a2 = x_step(cf, xm, mid, nxt, G2)
            xf += r[I_J2]
            assert close(a1, r[I_A1]) and close(a2, r[I_A2]) and close(xf, r[I_XF]), (k, xf, r[I_XF])
            if pd:
                sumf += (math.exp(xm) + math.exp(xf)) * dtf
            # coarse path: one plain step of 2 dt_f on the aggregated increments
            Gc = (a1 * G1 + a2 * G2) / math.hypot(a1, a2)
            assert close(Gc, r[I_GC])
            vcn = list(r[I_VC:I_VC + nf])
            if all_nv:
                if Z2 is not None:
                    want = [nv(K, v, (z1 + z2) / SQ2) for (_, K), v, z1, z2 in zip(cc, vc, Z1, Z2)]
                    assert all(close(a, b) for a, b in zip(want, vcn)), (k, want, vcn)
            else:
                assert vcn == nxt  # NCI / Exact: the fine variance, copied
            xc, _ = x_step(cc, xc, vc, vcn, Gc)
            xc += r[I_J1] + r[I_J2]
            assert close(xc, r[I_XC]), (k, xc, r[I_XC])
            if pd:
                sumc += math.exp(xc) * dtc
            vf, vc = nxt, vcn
        assert close(xf, math.log(ff[0])) and close(xc, math.log(fc[0]))
        if pd:
            # each path's tracker ran on its own grid
            assert close((sumf - 0.5 * ff[0] * dtf) / T, ff[4]) and close((sumc - 0.5 * fc[0] * dtc) / T, fc[4])


def test_aggregated_increments_are_independent_standard_normals(oracle):
    """Zc = (Z1 + Z2)/sqrt2 and Gc = (a1 G1 + a2 G2)/hypot(a1, a2) over 8,192
    recorded coarse steps: mean 0, variance 1, uncorrelated (4-sigma bands)."""
    m = make_model(1, [H1])
    inst = InstrumentSpec(strikes=[100.0], maturity=1.0)
    gc, zc = [], []
    for seed in range(256):
        rec, _, _, draws = oracle.debug_coupled_path(m, [capi.SCHEME_NV], inst, 32, 1000 + seed, S0, R, Q)
        for k in range(32):
            i0 = int(rec[k, 0])
            gc.append(rec[k, I_GC])
            zc.append((oracle.normcdfinv(draws[i0 + 2]) + oracle.normcdfinv(draws[i0 + 3])) / SQ2)
    gc, zc = np.array(gc), np.array(zc)
    n = gc.size
    for x in (gc, zc):
        assert abs(x.mean()) < 4 / math.sqrt(n)
        assert abs(x.var() - 1.0) < 4 * math.sqrt(2.0 / n)
    assert abs(np.corrcoef(gc, zc)[0, 1]) < 4 / math.sqrt(n)
