"""Host logic of the integrator registry (no GPU): names, refused
combinations, the stage fractions of the time blend."""
import numpy as np
import pytest

from kelp_amd import integrators as I
from kelp_amd import synth
from kelp_amd.diffusion import DiffusionStepArgs
from kelp_amd.fields import LazyLerpField, VelocityField
from kelp_amd.particles import make_state


def test_registry_names():
    assert I.get_integrator("rk2") is I.rk2_step
    # integrators.py:6635 INTEGRATORS
    assert set(I.INTEGRATORS) == {"euler", "rk2", "rk4", "rk45"}
    assert I.get_integrator("rk45") is I.rk45_step
    with pytest.raises(ValueError, match="Unknown integrator"):
        I.get_integrator("rk5")


def test_stage_alphas_match_reference_formula():
    # integrators.py:1187: alpha_k = c_k * dt / dt_to_next
    assert I.stage_alphas(4, 1800.0, 21600.0) == [0.0, 1800.0 * 0.5 / 21600.0,
                                                   1800.0 * 0.5 / 21600.0, 1800.0 / 21600.0]
    assert I.stage_alphas(2, 600.0, 600.0) == [0.0, 0.5, 0.0, 0.0]
    assert I.stage_alphas(1, 600.0, None) == [0.0] * 4
    with pytest.raises(ValueError, match="past the snapshot bracket"):
        I.stage_alphas(2, 700.0, 600.0)


@pytest.fixture(scope="module")
def small():
    mesh = synth.lattice_mesh(0.0, 4.0, 30.0, 34.0, 0.25, jitter=0.05, seed=1)
    u = np.zeros(mesh.n_nodes, np.float32)
    return mesh, VelocityField(u=u, v=u, uv_layout="node")


def _state():
    return make_state(np.array([1.0, 2.0]), np.array([31.0, 32.0]))


def test_refused_combinations_raise_before_the_device(small):
    mesh, vel = small
    st = _state()
    with pytest.raises(NotImplementedError, match="uv_interp"):
        I.rk2_step(st, vel, mesh, 600.0, 1.0, uv_interp="leastsq")
    with pytest.raises(NotImplementedError, match="uv_interp"):
        I.rk2_step(st, vel, mesh, 600.0, 1.0, uv_interp="edge_leastsq")
    el = VelocityField(u=np.zeros(mesh.n_elements), v=np.zeros(mesh.n_elements))
    with pytest.raises(NotImplementedError, match="uv_layout"):
        I.rk2_step(st, el, mesh, 600.0, 1.0, uv_interp="barycentric")
    d = DiffusionStepArgs("em", 1.0, 1.0, 1, 0, 0)
    with pytest.raises(NotImplementedError, match="scheme"):
        I.euler_step(st, vel, mesh, 600.0, 1.0, uv_interp="barycentric", diffusion=d)
    d = DiffusionStepArgs("uniform", 1.0, 1.0, 1, 0, 0, az_enabled=True, az_const=1e-4)
    with pytest.raises(ValueError, match="3-D"):
        I.euler_step(st, vel, mesh, 600.0, 1.0, uv_interp="barycentric", diffusion=d)
    d = DiffusionStepArgs("uniform", 1.0, 1.0, 1, 0, 0, az_enabled=True, az_requires_field=True)
    with pytest.raises(ValueError, match="no az field"):   # integrators.py:884
        I.euler_step(st, vel, mesh, 600.0, 1.0, uv_interp="barycentric", diffusion=d)
    with pytest.raises(ValueError, match="together"):
        I.rk4_step(st, vel, mesh, 600.0, 1.0, uv_interp="barycentric", vel_next=vel)
    with pytest.raises(ValueError, match="positive"):
        I.rk4_step(st, vel, mesh, -600.0, 1.0, uv_interp="barycentric")


def test_all_dead_returns_zero_velocity_without_device(small):
    # integrators.py:272 _zero_dead_velocity
    mesh, vel = small
    st = _state()
    st = st.replace(status=np.array([3, 2], np.int32), u=np.ones(2), v=np.ones(2))
    out = I.rk2_step(st, vel, mesh, 600.0, 1.0, uv_interp="barycentric")
    assert np.all(out.u == 0) and np.all(out.v == 0)
    assert np.all(st.u == 1)   # input untouched


def test_lazy_lerp_field_blends_like_reference():
    a = VelocityField(u=np.array([1.0, 2.0]), v=np.array([0.0, 4.0]), uv_layout="node")
    b = VelocityField(u=np.array([3.0, 2.0]), v=np.array([8.0, 0.0]), uv_layout="node")
    lz = LazyLerpField(a, b, 0.25)
    np.testing.assert_array_equal(lz.u, 0.75 * a.u + 0.25 * b.u)
    np.testing.assert_array_equal(lz.materialize().v, a.lerp(b, 0.25).v)
    with pytest.raises(ValueError):
        LazyLerpField(a, VelocityField(u=np.zeros(3), v=np.zeros(3), uv_layout="node"), 0.5)
