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


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_device_matches_reference_golden_pchip(gpu_ctx, direction):
    """4-snapshot PCHIP in time: the four snapshots resident in slots 3-6,
    each RK2 stage's field reconstructed on the device
    (kelp_field_reconstruct) into slots 0 and 2, then the step.  Also
    bit-identical to the kmath oracle doing the same in float64."""
    from oracle.oracle import OracleMesh, pchip4, step_args
    gold = np.load(os.path.join(GOLD, "fesom2_node_rk2_fused_pchip" + ("_bwd" if direction < 0 else "")
                                + ".npz"), allow_pickle=False)
    state = {}

    def stepper(mesh, st, snaps, times, tqs, i):
        if state.get("mesh") is not mesh:
            gpu_ctx.upload_mesh(mesh)
            for k, s in enumerate(snaps):
                gpu_ctx.upload_field(3 + k, s[0], s[1], s[2])
            state["mesh"] = mesh
            state["om"] = OracleMesh(mesh, "kmath")
            state["ost"] = dict(st)
        for dst, tq in zip((0, 2), tqs):
            gpu_ctx.reconstruct(dst, [3, 4, 5, 6], times, tq)
        p = gpu_ctx.particles(st["lon"], st["lat"], st["z"], status=st["status"], nn=st["nn"], fr=st["fr"])
        gpu_ctx.step(p, 2, 0, rc.NODE_DT, direction, slot_b=-1, slot_t=2, alpha0=0.0, alphas=[0.0, 1.0],
                     exact_nearest=2, w_method=2, step_i=i)
        out = p.to_host()
        R = [[pchip4(*(s[c] for s in snaps), times, tq, "kmath") for c in range(3)] for tq in tqs]
        fields = dict(u_a=R[0][0], v_a=R[0][1], w_a=R[0][2], u_t=R[1][0], v_t=R[1][1], w_t=R[1][2],
                      alpha0=0.0)
        want = state["om"].step(state["ost"], fields, step_args(2, rc.NODE_DT, direction, is3d=True,
                                                                w_method=2, exact=2, alphas=(0.0, 1.0),
                                                                step_i=i))
        state["ost"] = want
        for k in ("status", "fr", "nn"):
            np.testing.assert_array_equal(out[k], want[k], err_msg=f"{k} step {i}")
        for k in ("lon", "lat", "z", "u", "v", "w"):
            np.testing.assert_array_equal(out[k].view(np.int64), want[k].view(np.int64), err_msg=f"{k} step {i}")
        return out

    got = rc.run_pchip_case(direction, stepper)
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
