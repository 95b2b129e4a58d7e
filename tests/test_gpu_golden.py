"""The device step against the reference's own golden vectors (GPU).

The reference's Numba JIT outputs (tests/golden/reference_golden_3d) for the
node-layout HEALPix cases -- fused 3-D RK2 with a time blend forward and
backward, unfused 3-D RK4, the split anisotropic Kh walk forward and
backward, the vertical Az kick -- re-run through ``kelp_step`` on the
MI355X.  The dual mesh mixes quads and triangles, so this runs the
four-vertex kernels.  Bar: the reference's cross-environment rule
(tests/test_golden_3d_runs.py:54-72): integers exact, floats within
rtol = atol = 1e-9.
"""
import os

import numpy as np
import pytest

from .golden import reference_cases as rc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_golden_3d")


def _device_stepper(ctx):
    state = {}

    def step(mesh, st, fields, order, alphas, i, kick):
        if state.get("mesh") is not mesh:
            ctx.upload_mesh(mesh)
            state["mesh"] = mesh
        ctx.upload_field(0, fields["u_a"], fields["v_a"], fields["w_a"])
        ctx.upload_field(2, fields["u_t"], fields["v_t"], fields["w_t"])
        slot_b = -1
        if "u_b" in fields:
            ctx.upload_field(1, fields["u_b"], fields["v_b"], fields["w_b"])
            slot_b = 1
        p = ctx.particles(st["lon"], st["lat"], st["z"], status=st["status"], nn=st["nn"],
                          fr=st["fr"])
        ctx.step(p, order, 0, rc.NODE_DT, st.get("_dir", 1.0), slot_b=slot_b, slot_t=2,
                 alpha0=fields["alpha0"], alphas=alphas, exact_nearest=2, w_method=2,
                 kick=kick, step_i=i)
        out = p.to_host()
        out["_dir"] = st.get("_dir", 1.0)
        return out
    return step


@pytest.mark.parametrize("name,direction", [
    ("fesom2_node_rk2_fused_lerp", 1.0), ("fesom2_node_rk2_fused_lerp", -1.0),
    ("fesom2_node_rk4_unfused", 1.0),
    ("split_uniform_diffusion_rk2", 1.0), ("split_uniform_diffusion_rk2", -1.0),
    ("az_const_euler", 1.0),
])
def test_device_matches_reference_golden(gpu_ctx, name, direction):
    fname = name + ("_bwd" if direction < 0 else "") + ".npz"
    gold = np.load(os.path.join(GOLD, fname), allow_pickle=False)
    stepper = _device_stepper(gpu_ctx)

    def wrapped(mesh, st, fields, order, alphas, i, kick):
        return stepper(mesh, dict(st, _dir=direction), fields, order, alphas, i, kick)

    got = rc.run_case(name, direction, wrapped)
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


def _device_elem_stepper(ctx, direction):
    state = {}

    def step(mesh, st, fields, order, alphas, i):
        if state.get("mesh") is not mesh:
            ctx.upload_mesh(mesh)
            state["mesh"] = mesh
        ctx.upload_field(0, fields["u_a"], fields["v_a"], fields["w_a"], uv_layout="element")
        slot_b, slot_t = -1, -1
        if "u_t" in fields:
            ctx.upload_field(2, fields["u_t"], fields["v_t"], fields["w_t"], uv_layout="element")
            slot_t = 2
        if "u_b" in fields:
            ctx.upload_field(1, fields["u_b"], fields["v_b"], fields["w_b"], uv_layout="element")
            slot_b = 1
        p = ctx.particles(st["lon"], st["lat"], st["z"], status=st["status"], nn=st["nn"],
                          fr=st["fr"])
        ctx.step(p, order, 0, rc.ELEM_DT, direction, slot_b=slot_b, slot_t=slot_t,
                 alpha0=fields["alpha0"], alphas=alphas, exact_nearest=2, w_method=1, step_i=i)
        return p.to_host()
    return step


@pytest.mark.parametrize("name,direction", [
    ("fesom2_elem_euler_fused", 1.0),
    ("fesom2_elem_rk2_unfused_lerp", 1.0), ("fesom2_elem_rk2_unfused_lerp", -1.0),
])
def test_device_matches_reference_golden_element_layout(gpu_ctx, name, direction):
    """u/v at element centres by least squares, leastsq W."""
    fname = name + ("_bwd" if direction < 0 else "") + ".npz"
    gold = np.load(os.path.join(GOLD, fname), allow_pickle=False)
    got = rc.run_elem_case(name, direction, _device_elem_stepper(gpu_ctx, direction))
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
