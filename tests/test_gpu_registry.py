"""The integrator registry mirror (``get_integrator(name)(state, vel, mesh,
dt, time_dir, **kwargs)``, numerics/integrators.py:6635-6660) on the paths
beyond the plain z-level step: a reference-shaped ``Mesh`` (the attribute
names of src/kelp/core/mesh.py:41-210, grid hash in ``_gh_*``), sigma
meshes with the ``sigma_eta_start`` / ``sigma_eta_end`` keywords, a PCHIP
``recon`` bundle and the edge walk of a state carrying ``elem`` (GPU).

Bar: bit-identical to the oracle built with the device's math (kmath), or
to the same run through this package's own ``Mesh``; the PCHIP run also
meets the reference's golden vectors at rtol = atol = 1e-9.
"""
import dataclasses
import os
from typing import Any, Dict, Optional

import numpy as np
import pytest

from kelp_amd import integrators as I
from kelp_amd import synth
from kelp_amd.fields import LazyLerpField, VelocityField
from kelp_amd.mesh import build_elem_neighbors, from_reference_mesh, set_exact_nearest
from kelp_amd.particles import make_state
from oracle.oracle import OracleMesh, _blend, sigma_stage_etas, stage_field, step_args

from .golden import reference_cases as rc

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float64).view(np.int64)


@dataclasses.dataclass
class RefMesh:
    """The reference Mesh's fields (core/mesh.py:81-158), by name."""
    name: str
    nodes_lon: np.ndarray
    nodes_lat: np.ndarray
    depth: np.ndarray
    elem_nodes: np.ndarray
    elem_center_lon: np.ndarray
    elem_center_lat: np.ndarray
    node_to_elems: np.ndarray
    node_to_elems_count: np.ndarray
    sigma: np.ndarray = dataclasses.field(default_factory=lambda: np.array([], dtype=np.float64))
    vertical_coord: str = "zlevel"
    r_earth: float = 6_371_000.0
    metadata: Dict[str, Any] = dataclasses.field(default_factory=dict)
    _gh_offsets: Optional[np.ndarray] = None
    _gh_node_ids: Optional[np.ndarray] = None
    _gh_cell_size: float = 0.0
    _gh_lon_min: float = 0.0
    _gh_lat_min: float = 0.0
    _gh_nx: int = 0
    _gh_ny: int = 0
    _gh_wrap_lon: bool = False
    _gh_far_rows: int = 16
    _gh_max_edge_deg: Optional[float] = None
    _elem_is_coastal: Optional[np.ndarray] = None
    _elem_neighbors: Optional[np.ndarray] = None


def _as_reference(m) -> RefMesh:
    return RefMesh(m.name, m.nodes_lon, m.nodes_lat, m.depth, m.elem_nodes, m.elem_center_lon,
                   m.elem_center_lat, m.node_to_elems, m.node_to_elems_count, m.sigma.copy(),
                   m.vertical_coord, m.r_earth, dict(m.metadata), m.gh_offsets, m.gh_node_ids,
                   m.gh_cell_size, m.gh_lon_min, m.gh_lat_min, m.gh_nx, m.gh_ny, m.gh_wrap_lon,
                   m.gh_far_rows, m.gh_max_edge_deg)


def test_reference_mesh_through_registry_bitwise():
    """rk4_step on a reference-shaped Mesh equals the run on this package's
    Mesh bit for bit; the adapter shares the arrays, is built once, and the
    device mesh is uploaded once for all the steps."""
    mesh = synth.sphere_mesh(20000)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(2)]
    vels = [VelocityField(u=s[0], v=s[1], w=s[2], uv_layout="node") for s in snaps]
    lon, lat, z = synth.wet_area_seeds(mesh, 3000)
    ref = _as_reference(mesh)
    ref.metadata["exact_nearest"] = True
    set_exact_nearest(mesh, True)
    adapted = from_reference_mesh(ref)
    assert adapted is from_reference_mesh(ref)
    assert adapted.nodes_lon is ref.nodes_lon and adapted.gh_offsets is ref._gh_offsets
    assert adapted.metadata is ref.metadata
    eng = I.get_engine()
    uploads = []
    orig = eng.ctx.upload_mesh
    eng.ctx.upload_mesh = lambda m: (uploads.append(m), orig(m))[1]
    try:
        step = I.get_integrator("rk4")
        runs = {}
        for key, msh in (("ref", ref), ("own", mesh)):
            st = make_state(lon, lat, z)
            for k in range(6):
                vel = LazyLerpField(vels[0], vels[1], 0.1 * k)
                st = step(st, vel, msh, 1800.0, 1.0, uv_interp="barycentric", w_interp="barycentric")
            runs[key] = st
    finally:
        eng.ctx.upload_mesh = orig
    assert len(uploads) == 2            # one per mesh object, not one per step
    a, b = runs["ref"], runs["own"]
    for k in ("lon", "lat", "z", "u", "v", "w"):
        assert np.array_equal(_bits(getattr(a, k)), _bits(getattr(b, k))), k
    for k in ("status", "nearest_node", "failure_reason"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert (a.status == 1).mean() > 0.9


@pytest.fixture(scope="module")
def sigma_world():
    from .test_gpu_sigma import _sigma_mesh
    mesh = _sigma_mesh()
    rng = np.random.default_rng(31)
    K, N = mesh.sigma.size, mesh.n_nodes
    snaps = []
    for s in range(3):
        eta = 0.3 * np.sin(0.9 * s + mesh.nodes_lon) * np.cos(mesh.nodes_lat)
        u = 0.3 + 0.1 * rng.standard_normal((K - 1, N))
        v = -0.1 + 0.1 * rng.standard_normal((K - 1, N))
        w = 2e-4 * rng.standard_normal((K, N))
        w[0] = w[-1] = 0.0
        snaps.append(dict(eta=eta, u=u, v=v, w=w))
    P = 2000
    lon = rng.uniform(0.1, 2.9, P)
    lat = rng.uniform(40.1, 42.4, P)
    H0 = 8.0 + 12.0 * (lon / 3.0) + 5.0 * np.sin(lat * 2.0)
    z = rng.uniform(0.0, 1.0, P) * np.maximum(H0, 0.5)
    return mesh, snaps, lon, lat, z


@pytest.mark.parametrize("given", [True, False])
@pytest.mark.parametrize("order,w_interp", [(2, "idw"), (4, "barycentric"), (1, "leastsq")])
def test_sigma_through_registry_bitwise(sigma_world, order, w_interp, given):
    """A time-blended sigma step through the registry: the surfaces given as
    sigma_eta_start / sigma_eta_end arrays, or left to the step (the field's
    own eta at the start, its blend towards vel_next at the end)."""
    mesh, snaps, lon, lat, z = sigma_world
    vels = [VelocityField(u=s["u"], v=s["v"], w=s["w"], eta=s["eta"], uv_layout="node") for s in snaps]
    a0, dtn, dt = 0.3, 3000.0, 600.0
    vel = LazyLerpField(vels[0], vels[1], a0)
    eta = [s["eta"] for s in snaps]
    kw = dict(uv_interp="barycentric", w_interp=w_interp, vel_next=vels[2], dt_to_next=dtn,
              Dmin=0.5, z0b_min=0.1)
    if given:
        e0, ee = _blend(eta[0], eta[1], 0.25), _blend(eta[1], eta[2], 0.6)
        kw.update(sigma_eta_start=e0, sigma_eta_end=ee)
    else:
        e0 = stage_field(eta[0], eta[1], eta[2], a0, 0.0)
        ee = stage_field(eta[0], eta[1], eta[2], a0, dt / dtn)
    st = I.get_integrator({1: "euler", 2: "rk2", 4: "rk4"}[order])(make_state(lon, lat, z), vel, mesh,
                                                                     dt, 1.0, **kw)
    al = I.stage_alphas(order, dt, dtn)
    etas = sigma_stage_etas(order, False, e0, ee, eta[0], eta[1], eta[2], a0, al)
    f = dict(u_a=snaps[0]["u"], v_a=snaps[0]["v"], w_a=snaps[0]["w"], u_b=snaps[1]["u"],
             v_b=snaps[1]["v"], w_b=snaps[1]["w"], u_t=snaps[2]["u"], v_t=snaps[2]["v"],
             w_t=snaps[2]["w"], alpha0=a0, uv_layout="node")
    wm = I.SIGMA_W_METHODS[w_interp]
    args = step_args(order, dt, 1.0, is3d=True, w_method=wm, exact=2, alphas=al,
                     sigma=dict(etas=etas, Dmin=0.5, z0b_min=0.1))
    om = OracleMesh(mesh, "kmath")
    want = om.step(dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
                        nn=np.full(lon.size, -1, np.int32), pid=np.arange(lon.size)), f, args)
    assert (want["status"] == 1).mean() > 0.8
    np.testing.assert_array_equal(st.status, want["status"])
    np.testing.assert_array_equal(st.nearest_node, want["nn"])
    np.testing.assert_array_equal(st.failure_reason, want["fr"])
    for k in ("lon", "lat", "z", "u", "v", "w"):
        assert np.array_equal(_bits(getattr(st, k)), _bits(want[k])), k


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_pchip_recon_through_registry_matches_reference_golden(direction):
    """rk2_step with the reference's recon bundle (fields, times, q,
    direction) on the golden case fesom2_node_rk2_fused_pchip."""
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_golden_3d",
                                "fesom2_node_rk2_fused_pchip" + ("_bwd" if direction < 0 else "") + ".npz"),
                   allow_pickle=False)
    cache = {}

    def stepper(mesh, st, snaps, times, tqs, i):
        if cache.get("mesh") is not mesh:
            cache["mesh"] = mesh
            cache["vels"] = [VelocityField(u=s[0], v=s[1], w=s[2], uv_layout="node") for s in snaps]
            cache["st"] = make_state(st["lon"], st["lat"], st["z"])
        q = tqs[0]
        out = I.rk2_step(cache["st"], cache["vels"][0], mesh, rc.NODE_DT, direction,
                         uv_interp="barycentric", w_interp="barycentric",
                         recon=(cache["vels"], times, q, direction))
        cache["st"] = out
        return dict(lon=out.lon, lat=out.lat, z=out.z, status=out.status, fr=out.failure_reason,
                    u=out.u, v=out.v, w=out.w, nn=out.nearest_node)

    got = rc.run_pchip_case(direction, stepper)
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


def test_walk_through_registry_bitwise():
    """A state carrying ``elem`` runs the edge walk: rk2_step returns the new
    carried elements; bit-identical to the oracle's walk over 6 steps."""
    mesh = synth.sphere_mesh(12000)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(2)]
    vels = [VelocityField(u=s[0].astype(np.float32), v=s[1].astype(np.float32),
                          w=s[2].astype(np.float32), uv_layout="node") for s in snaps]
    lon, lat, z = synth.wet_area_seeds(mesh, 3000)
    om = OracleMesh(mesh, "kmath")
    nn0 = om.nearest(lon, lat, exact=True)
    elem0 = om.containing(lon, lat, nn0)
    set_exact_nearest(mesh, True)
    st = make_state(lon, lat, z).replace(elem=elem0.astype(np.int32))
    want = dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
                nn=np.full(lon.size, -1, np.int32), pid=np.arange(lon.size), elem=elem0.copy())
    build_elem_neighbors(mesh)
    om = OracleMesh(mesh, "kmath")
    dt = 1800.0
    for k in range(6):
        a0 = 0.15 * k
        st = I.rk2_step(st, LazyLerpField(vels[0], vels[1], a0), mesh, dt, 1.0,
                        uv_interp="barycentric", w_interp="barycentric")
        f = dict(u_a=vels[0].u, v_a=vels[0].v, w_a=vels[0].w, u_b=vels[1].u, v_b=vels[1].v,
                 w_b=vels[1].w, alpha0=a0)
        want = om.step(want, f, step_args(2, dt, 1.0, is3d=True, w_method=2, exact=1, walk=True))
    np.testing.assert_array_equal(st.elem, want["elem"])
    np.testing.assert_array_equal(st.status, want["status"])
    np.testing.assert_array_equal(st.nearest_node, want["nn"])
    for k in ("lon", "lat", "z", "u", "v", "w"):
        assert np.array_equal(_bits(getattr(st, k)), _bits(want[k])), k
