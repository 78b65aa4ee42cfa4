"""Pins the oracle's F-bar projection (fbar_projection.cpp:18-94) by what
the mathematics fixes; CPU only.

* the nodal projection conserves the volume-weighted J:
  sum_a nJbar_a nvof_a = sum_p Jbar_p det(dF_p) V_p (a node-mass weight, the
  initial volume or a dropped division fails);
* a constant J field is reproduced (Jbar = J) and then leaves dF unscaled;
* a homogeneous deformation (affine velocity carried by TPIC) gives
  Fbar = F and Jbar = J;
* det(Fbar) = Jbar holds along a run (the cbrt scaling and the product
  order of Fbar = dF Fbar_old)."""
import numpy as np
import pytest

import pyoracle
from mpm_mi355x import capi, scenes

from test_oracle_affine import A, filled_scene


def prepared(idimc=(1, 1, 1), seed=3):
    s = filled_scene(idimc, pyoracle.PIC)
    rng = np.random.default_rng(seed)
    s.parts.vol = s.parts.vol0 * (0.5 + rng.random(s.n))  # V != V0
    s.fbar = True
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    o.p2g()
    return o, rng


@pytest.mark.parametrize("idimc", [(1, 1, 1), (2, 1, 2)])
def test_projection_conserves_volume_weighted_J(idimc):
    o, rng = prepared(idimc)
    n = o.parts.n
    o.det_def_grad_bar[:] = 0.8 + 0.4 * rng.random(n)
    o.det_ddg[:] = 0.9 + 0.2 * rng.random(n)
    o.ddg[:] = np.tile(np.eye(3).ravel(), n)
    want = np.sum(o.det_def_grad_bar * o.det_ddg * o.parts.vol)
    o.fbar()
    has = o.nvof > o.st.mtol
    got = np.sum(o.njbar[has] * o.nvof[has])
    assert abs(got - want) <= 1e-12 * want
    assert np.all(o.njbar[~has] == 0.0)


def test_constant_J_is_reproduced_and_leaves_dF_unscaled():
    o, rng = prepared((2, 1, 2))
    n = o.parts.n
    o.det_def_grad_bar[:] = 1.25
    o.det_ddg[:] = 0.8       # J = 1 for every particle
    dd = np.eye(3) + 1e-3 * rng.normal(size=(n, 3, 3))
    o.ddg[:] = dd.ravel()
    o.fbar()
    has = o.nvof > o.st.mtol
    np.testing.assert_allclose(o.njbar[has], 1.0, rtol=1e-14)
    np.testing.assert_allclose(o.det_def_grad_bar, 1.0, rtol=1e-14)
    np.testing.assert_allclose(o.ddg.reshape(n, 3, 3), dd, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(o.def_grad_bar, dd, rtol=1e-14, atol=1e-16)   # Fbar_old = I


def test_homogeneous_deformation_gives_Fbar_equal_F():
    s = filled_scene((2, 2, 2), pyoracle.TPIC)
    s.parts.vel[:] = s.parts.coord @ (0.5 * A).T
    s.fbar = True
    o = pyoracle.Oracle(s)
    o.cmat[:] = 0.5 * A
    assert o.step(1) == 0
    F = o.parts.def_grad
    np.testing.assert_allclose(o.def_grad_bar, F, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.det_def_grad_bar, o.parts.det_def_grad, rtol=1e-12)
    assert np.abs(F - np.eye(3)).max() > 1e-6


@pytest.mark.parametrize("model", [capi.MODEL_NEOHOOKEAN, capi.MODEL_LINELAST])
def test_det_Fbar_equals_Jbar_along_a_run(model):
    s = scenes.tiny(n_cells=8, model=model, props=(5e5, 0.35))
    s.fbar = True
    o = pyoracle.Oracle(s)
    assert o.run(1, 60) == 0
    np.testing.assert_allclose(np.linalg.det(o.def_grad_bar), o.det_def_grad_bar, rtol=1e-10)
    # the projection smooths J: Jbar differs from J, within the range of J
    J = o.parts.det_def_grad
    assert np.abs(o.det_def_grad_bar - J).max() > 1e-9
    assert o.det_def_grad_bar.min() >= J.min() - 1e-6 and o.det_def_grad_bar.max() <= J.max() + 1e-6
