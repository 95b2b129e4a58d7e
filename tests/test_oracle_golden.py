"""The oracle against the reference's Numba golden vectors (CPU).

Pins the 3-D z-level path of the oracle end to end -- HEALPix dual mesh with
quads and triangles, layer search, barycentric/Wachspress u/v, barycentric W,
linear time blend, fused RK2 / unfused RK4, the split anisotropic Kh kick with
its counter RNG -- against outputs of the reference's own JIT kernels
(tests/golden/reference_golden_3d, copied from the reference).  Bar: the
reference's own cross-environment rule (tests/test_golden_3d_runs.py:54-72):
integers exact, floats within rtol = atol = 1e-9.
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleMesh, step_args

from .golden import reference_cases as rc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "reference_golden_3d")

_MESHES = {}


def _oracle_stepper(flavour):
    def step(mesh, st, fields, order, alphas, i, kick):
        om = _MESHES.get((id(mesh), flavour))
        if om is None:
            om = _MESHES[(id(mesh), flavour)] = OracleMesh(mesh, flavour)
        # guard off (meshes built outside run_tracking): the batch 3x3 anchor
        args = step_args(order, rc.NODE_DT, st.get("_dir", 1.0), is3d=True, w_method=2,
                         exact=2, alphas=alphas, kick=kick, step_i=i)
        out = om.step(st, fields, args)
        out["_dir"] = st.get("_dir", 1.0)
        return out
    return step


@pytest.mark.parametrize("name,direction", [
    ("fesom2_node_rk2_fused_lerp", 1.0), ("fesom2_node_rk2_fused_lerp", -1.0),
    ("fesom2_node_rk4_unfused", 1.0),
    ("split_uniform_diffusion_rk2", 1.0), ("split_uniform_diffusion_rk2", -1.0),
    ("az_const_euler", 1.0),
])
@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_oracle_matches_reference_golden(name, direction, flavour):
    fname = name + ("_bwd" if direction < 0 else "") + ".npz"
    gold = np.load(os.path.join(GOLD, fname), allow_pickle=False)
    stepper = _oracle_stepper(flavour)

    def wrapped(mesh, st, fields, order, alphas, i, kick):
        st = dict(st, _dir=direction)
        return stepper(mesh, st, fields, order, alphas, i, kick)

    got = rc.run_case(name, direction, wrapped)
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


@pytest.mark.parametrize("name,direction", [
    ("fesom2_elem_euler_fused", 1.0),
    ("fesom2_elem_rk2_unfused_lerp", 1.0), ("fesom2_elem_rk2_unfused_lerp", -1.0),
])
@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_oracle_matches_reference_golden_element_layout(name, direction, flavour):
    """Element-centre u/v by least squares and leastsq W (steady fused Euler,
    time-blended unfused RK2 forward and backward)."""
    fname = name + ("_bwd" if direction < 0 else "") + ".npz"
    gold = np.load(os.path.join(GOLD, fname), allow_pickle=False)

    def stepper(mesh, st, fields, order, alphas, i):
        om = _MESHES.get((id(mesh), flavour))
        if om is None:
            om = _MESHES[(id(mesh), flavour)] = OracleMesh(mesh, flavour)
        args = step_args(order, rc.ELEM_DT, direction, is3d=True, w_method=1, exact=2,
                         alphas=alphas, step_i=i)
        return om.step(st, fields, args)

    got = rc.run_elem_case(name, direction, stepper)
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


@pytest.mark.parametrize("direction", [1.0, -1.0])
@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_oracle_matches_reference_golden_pchip(direction, flavour):
    """4-snapshot PCHIP in time (the fused 3-D barycentric RK2 PCHIP kernel,
    integrators.py:3859): the stage fields reconstructed by the oracle's
    _pchip_value_scalar restatement, then the RK2 step on them."""
    from oracle.oracle import pchip4
    gold = np.load(os.path.join(GOLD, "fesom2_node_rk2_fused_pchip" + ("_bwd" if direction < 0 else "")
                                + ".npz"), allow_pickle=False)

    def stepper(mesh, st, snaps, times, tqs, i):
        om = _MESHES.get((id(mesh), flavour))
        if om is None:
            om = _MESHES[(id(mesh), flavour)] = OracleMesh(mesh, flavour)
        R = [[pchip4(*(s[c] for s in snaps), times, tq, flavour) for c in range(3)] for tq in tqs]
        fields = dict(u_a=R[0][0], v_a=R[0][1], w_a=R[0][2], u_t=R[1][0], v_t=R[1][1], w_t=R[1][2],
                      alpha0=0.0)
        args = step_args(2, rc.NODE_DT, direction, is3d=True, w_method=2, exact=2,
                         alphas=(0.0, 1.0), step_i=i)
        return om.step(st, fields, args)

    got = rc.run_pchip_case(direction, stepper)
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
