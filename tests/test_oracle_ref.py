"""The FP64 oracle against the reference's own numerics, bit for bit.

oracle/_ref/libmpm_ref.so is compiled (oracle/Makefile) from the reference's
module/solid/constitutive_model.cpp and module/bc.cpp plus a harness that
instantiates module/cal_mat.h and module/map_and_interpolate.h. When it is
not built (no reference tree), these tests are skipped.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle
from pyoracle import P

ref = pyoracle.load_ref()
orc = pyoracle.load()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (reference tree absent)")


def rand_F(rng, scale=0.2):
    F = np.eye(3) + scale * rng.standard_normal((3, 3))
    if np.linalg.det(F) <= 0:
        F[0] *= -1
    return F


MODELS = [
    (-1, [1e6, 0.3]),
    (0, [1e6, 0.3]),
    (1, [5e5, 0.35]),
    (2, [2e6, 0.25]),
    (3, [1.5e5, 0.4e5, 1e7]),
]


@needs_ref
@pytest.mark.parametrize("model,props", MODELS)
def test_update_stress_bitexact(model, props):
    rng = np.random.default_rng(1234 + model)
    for _ in range(200):
        F = rand_F(rng)
        dF = np.eye(3) + 1e-3 * rng.standard_normal((3, 3))
        s0 = rng.standard_normal(6) * 1e3
        jac = float(np.linalg.det(F))
        jbar = jac if rng.random() < 0.7 else jac * (1 + 1e-3 * rng.standard_normal())
        pr = np.zeros(8)
        pr[: len(props)] = props
        a = s0.copy()
        b = s0.copy()
        Fc, dFc = np.ascontiguousarray(F), np.ascontiguousarray(dF)
        ref.ref_update_stress(model, P(a), P(Fc), P(dFc), P(pr), len(props), jac, jbar)
        orc.orc_update_stress(model, P(b), P(Fc), P(dFc), P(pr), jac, jbar)
        assert np.array_equal(a, b), (model, a, b)


@needs_ref
@pytest.mark.parametrize("op", [0, 1, 2])
def test_bc_ops_bitexact(op):
    """BCSetVal / BCSetZero against orc_bc_set_val / orc_bc_set_zero; BCSetDt
    (fbc * dt * facl) against its closed form. Lists contain duplicates, so the
    later-entry-wins order is checked too."""
    rng = np.random.default_rng(7 + op)
    nodec = 50
    dt = 1e-3
    for facl in (1.0, 0.37, 2.5):
        nbc = rng.integers(0, nodec, size=30).astype(np.int32)
        fbc = rng.standard_normal(30)
        base = rng.standard_normal(3 * nodec)
        a = base.copy()
        ref.ref_bc_apply(op, len(nbc), P(nbc, C.c_int32), P(fbc), nodec, P(a), len(a), facl, dt)
        b = base.copy()
        st = pyoracle.OrcState()
        st.ibc[0] = len(nbc)
        st.nbc[0] = P(nbc, C.c_int32)
        st.fbc[0] = P(fbc)
        if op == 0:
            orc.orc_bc_set_val(C.byref(st), 0, nodec, P(b), facl)
        elif op == 1:
            orc.orc_bc_set_zero(C.byref(st), 0, nodec, P(b))
        else:
            for i in range(len(nbc)):
                b[nbc[i] + nodec] = fbc[i] * dt * facl
        assert np.array_equal(a, b)


@needs_ref
def test_det_inv_bitexact():
    rng = np.random.default_rng(99)
    for _ in range(500):
        A = np.ascontiguousarray(rng.standard_normal((3, 3)))
        assert ref.ref_det3(P(A)) == orc.orc_det3(P(A))
        o1, o2 = np.zeros(9), np.zeros(9)
        ref.ref_inv3(P(A), P(o1))
        orc.orc_inv3(P(A), P(o2))
        assert np.array_equal(o1, o2)


@needs_ref
@pytest.mark.parametrize("scheme", [0, 1])
def test_var_g2p_accumulation_order(scheme):
    """PIC/FLIP::VarG2P accumulate (sfi*node) resp. (sfi*dt*node) in pair order;
    the oracle's G2P uses the same expressions (checked on a pair list)."""
    rng = np.random.default_rng(5 + scheme)
    n, nodec, npairs = 7, 20, 60
    pv = rng.standard_normal((n, 3))
    node = rng.standard_normal(3 * nodec)
    pid = rng.integers(0, n, npairs).astype(np.int32)
    nid = rng.integers(0, nodec, npairs).astype(np.int32)
    sfi = rng.random(npairs)
    dt = 3.7e-4
    a = np.ascontiguousarray(pv.copy())
    ref.ref_var_g2p(scheme, n, P(a), nodec, P(node), npairs, P(pid, C.c_int32), P(nid, C.c_int32), P(sfi), dt)
    b = pv.copy()
    for k in range(npairs):
        for c in range(3):
            if scheme == 1:
                b[pid[k], c] += sfi[k] * dt * node[nid[k] + c * nodec]
            else:
                b[pid[k], c] += sfi[k] * node[nid[k] + c * nodec]
    assert np.array_equal(a, b)
