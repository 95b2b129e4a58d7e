"""Device sigma-coordinate step against the oracle (GPU).

A FESOM-C-like mesh (a 0.15 deg lattice of triangles with a band of quads,
11 stretched sigma interfaces, sloping bed with a shallow bank) with three
snapshots of u/v (at the nodes or at the element centres), the dia-sigma
flux W and a surface elevation each.  Every integrator, every W / column
height method (nearest, leastsq, barycentric, idw), steady steps and time-
blended ones with supplied start/end surfaces and with the end surface left
to the stepper, forward and backward.  Bar: bit-identical to the oracle
built with the device's math (kmath) in every output; the glibc oracle to
1e-9 deg / 1e-6 m.
"""
import numpy as np
import pytest

from kelp_amd.mesh import build_grid_hash, make_mesh
from oracle.oracle import OracleMesh, _blend, sigma_stage_etas, stage_field, step_args

pytestmark = pytest.mark.gpu

SIG = 1.0 - np.linspace(0.0, 1.0, 11) ** 2
DT = 600.0


def _sigma_mesh():
    xs = np.arange(0.0, 3.0 + 1e-9, 0.15)
    ys = np.arange(40.0, 42.5 + 1e-9, 0.15)
    nx, ny = xs.size, ys.size
    gx, gy = np.meshgrid(xs, ys)
    lon, lat = gx.ravel(), gy.ravel()
    elems = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            n0 = j * nx + i
            q = [n0, n0 + 1, n0 + nx + 1, n0 + nx]
            if 6 <= j <= 9:
                elems.append(q)
            else:
                elems += [[q[0], q[1], q[2], -1], [q[0], q[2], q[3], -1]]
    depth = 8.0 + 12.0 * (lon / 3.0) + 5.0 * np.sin(lat * 2.0)
    depth[(np.abs(lon - 1.5) < 0.2) & (np.abs(lat - 41.2) < 0.2)] = 0.6   # a bank drying at low tide
    mesh = make_mesh("sigma_lattice", lon, lat, np.array(elems, np.int32), depth=depth,
                     r_earth=6_400_000.0)
    mesh.sigma = SIG.copy()
    mesh.vertical_coord = "sigma"
    build_grid_hash(mesh)
    return mesh


@pytest.fixture(scope="module")
def world():
    mesh = _sigma_mesh()
    rng = np.random.default_rng(11)
    K, N, E = SIG.size, mesh.n_nodes, mesh.n_elements
    snaps = []
    for s in range(3):
        eta = 0.4 * np.sin(0.7 * s + mesh.nodes_lon) * np.cos(mesh.nodes_lat) + 0.05 * rng.standard_normal(N)
        un = 0.3 + 0.1 * rng.standard_normal((K - 1, N))
        vn = -0.1 + 0.1 * rng.standard_normal((K - 1, N))
        ue = 0.3 + 0.1 * rng.standard_normal((K - 1, E))
        ve = -0.1 + 0.1 * rng.standard_normal((K - 1, E))
        w = 2e-4 * rng.standard_normal((K, N))
        w[0] = 0.0
        w[-1] = 0.0
        snaps.append(dict(eta=eta, un=un, vn=vn, ue=ue, ve=ve, w=w))
    P = 3000
    lon = rng.uniform(0.1, 2.9, P)
    lat = rng.uniform(40.1, 42.4, P)
    H0 = 8.0 + 12.0 * (lon / 3.0) + 5.0 * np.sin(lat * 2.0)
    z = rng.uniform(0.0, 1.0, P) * np.maximum(H0, 0.5)
    pid = np.arange(P, dtype=np.int64)
    return mesh, snaps, lon, lat, z, pid


@pytest.mark.parametrize("layout", ["node", "element"])
@pytest.mark.parametrize("w_method", [0, 1, 2, 3])
@pytest.mark.parametrize("order", [1, 2, 4])
def test_sigma_step_bitwise(gpu_ctx, world, order, w_method, layout):
    mesh, snaps, lon, lat, z0, pid = world
    gpu_ctx.upload_mesh(mesh)
    uk, vk = ("un", "vn") if layout == "node" else ("ue", "ve")
    for s, sn in enumerate(snaps):
        gpu_ctx.upload_field(s, sn[uk], sn[vk], sn["w"], uv_layout=layout)
        gpu_ctx.upload_eta(s, sn["eta"])
    om = OracleMesh(mesh, "kmath")
    cases = [  # (steady, alpha0, stage alphas, supply the end surface, direction)
        (True, 0.0, (0.0, 0.0, 0.0, 0.0), True, 1.0),
        (True, 0.0, (0.0, 0.0, 0.0, 0.0), False, -1.0),
        (False, 0.3, (0.0, 0.2, 0.2, 0.4), True, 1.0),
        (False, 0.3, (0.0, 0.2, 0.2, 0.4), False, -1.0),
    ]
    for steady, a0, alphas, supply_end, direction in cases:
        p = gpu_ctx.particles(lon, lat, z0, pid=pid)
        start = (0, 1, -1, 0.25, 0.0)
        end = (1, 2, -1, 0.6, 0.0) if supply_end else None
        sb, st = (-1, -1) if steady else (1, 2)
        a0_ = 0.0 if steady else a0
        al = [0.0] * 4 if steady else list(alphas[:order]) + [0.0] * (4 - order)
        if end is None and not steady:
            end = (0, 1, 2, a0_, al[3] if order == 4 else 2.0 * al[1] if order == 2 else 0.5)
        # the oracle's surfaces from the same specs
        eta = [sn["eta"] for sn in snaps]
        e0 = _blend(eta[0], eta[1], 0.25)
        if end is None:
            ee = None
        elif end[2] < 0:
            ee = _blend(eta[end[0]], eta[end[1]], end[3])
        else:
            ee = stage_field(eta[0], eta[1], eta[2], end[3], end[4])
        etas = sigma_stage_etas(order, steady, e0, ee, eta[0], None if steady else eta[1],
                                None if steady else eta[2], a0_, al)
        gpu_ctx.step(p, order, 0, DT, direction, slot_b=sb, slot_t=st, alpha0=a0_, alphas=al,
                     exact_nearest=2, w_method=w_method,
                     sigma=dict(start=start, end=end, Dmin=0.5, z0b_min=0.1))
        got = p.to_host()
        fa = snaps[0]
        flds = dict(u_a=fa[uk], v_a=fa[vk], w_a=fa["w"], alpha0=a0_, uv_layout=layout)
        if not steady:
            flds.update(u_b=snaps[1][uk], v_b=snaps[1][vk], w_b=snaps[1]["w"],
                        u_t=snaps[2][uk], v_t=snaps[2][vk], w_t=snaps[2]["w"])
        args = step_args(order, DT, direction, is3d=True, w_method=w_method, exact=2, alphas=al,
                         sigma=dict(etas=etas, Dmin=0.5, z0b_min=0.1))
        want = om.step(dict(lon=lon, lat=lat, z=z0, status=np.ones(lon.size, np.int32),
                            nn=np.full(lon.size, -1, np.int32), pid=pid), flds, args)
        tag = (order, w_method, layout, steady, supply_end, direction)
        assert (want["status"] == 1).mean() > 0.8, tag
        for k in ("status", "fr", "nn"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=f"{k} {tag}")
        for k in ("lon", "lat", "z", "u", "v", "w"):
            np.testing.assert_array_equal(got[k].view(np.int64), want[k].view(np.int64),
                                          err_msg=f"{k} {tag}")


@pytest.mark.parametrize("order,w_method,layout", [(2, 1, "element"), (4, 2, "node"), (1, 3, "element")])
def test_sigma_kicks_and_beaching_bitwise(gpu_ctx, world, order, w_method, layout):
    """The split Kh kick (sigma fraction kept in the end surface), the Az
    kick folded into the end column and depth/on-failure beaching with the
    surface, on time-blended sigma steps, three steps in a row."""
    mesh, snaps, lon, lat, z0, pid = world
    gpu_ctx.upload_mesh(mesh)
    uk, vk = ("un", "vn") if layout == "node" else ("ue", "ve")
    for s, sn in enumerate(snaps):
        gpu_ctx.upload_field(s, sn[uk], sn[vk], sn["w"], uv_layout=layout)
        gpu_ctx.upload_eta(s, sn["eta"])
    om = OracleMesh(mesh, "kmath")
    kick = dict(kh_x=20.0, kh_y=8.0, seed=123, az_const=5e-4)
    beach = dict(criteria=["depth", "on_failure"], depth_min=1.2, use_eta=True, resuspension="auto",
                 min_wet_steps=1, dt=DT)
    p = gpu_ctx.particles(lon, lat, z0, pid=pid)
    st = dict(lon=lon, lat=lat, z=z0, status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), pid=pid)
    eta = [sn["eta"] for sn in snaps]
    a0 = 0.3
    al = [0.0, 0.2, 0.2, 0.4][:order] + [0.0] * (4 - order)
    for k in range(3):
        start = (0, 1, -1, 0.25, 0.0)
        end = (1, 2, -1, 0.6, 0.0)
        gpu_ctx.step(p, order, 0, DT, 1.0, slot_b=1, slot_t=2, alpha0=a0, alphas=al,
                     exact_nearest=2, w_method=w_method, kick=kick, beach=beach, step_i=k,
                     sigma=dict(start=start, end=end, Dmin=0.5, z0b_min=0.1))
        etas = sigma_stage_etas(order, False, _blend(eta[0], eta[1], 0.25), _blend(eta[1], eta[2], 0.6),
                                eta[0], eta[1], eta[2], a0, al)
        flds = dict(u_a=snaps[0][uk], v_a=snaps[0][vk], w_a=snaps[0]["w"], u_b=snaps[1][uk],
                    v_b=snaps[1][vk], w_b=snaps[1]["w"], u_t=snaps[2][uk], v_t=snaps[2][vk],
                    w_t=snaps[2]["w"], alpha0=a0, uv_layout=layout, eta_a=eta[0], eta_b=eta[1])
        args = step_args(order, DT, 1.0, is3d=True, w_method=w_method, exact=2, alphas=al, kick=kick,
                         beach=beach, step_i=k, sigma=dict(etas=etas, Dmin=0.5, z0b_min=0.1))
        st = om.step(st, flds, args)
    got = p.to_host()
    assert (st["status"] == 2).any() and (st["status"] == 1).mean() > 0.5
    for key in ("status", "fr", "nn", "ws"):
        np.testing.assert_array_equal(got[key], st[key], err_msg=key)
    for key in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_array_equal(got[key].view(np.int64), st[key].view(np.int64), err_msg=key)
