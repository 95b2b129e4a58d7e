"""Record-point sampling on the device against the oracle (GPU).

Every method (nearest, leastsq, barycentric, idw) on a 2-D, a mid-layer
and an interface node scalar, plain and blended between two snapshots,
with non-finite source values (filtered by the 1-ring combinations) and
dead particles (NaN).  Bar: bit-identical to the oracle built with the
device's math, NaN where the oracle has NaN.
"""
import numpy as np
import pytest

from kelp_amd import synth
from oracle.oracle import OracleMesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    mesh = synth.sphere_mesh(6000)
    lon, lat, z = synth.wet_area_seeds(mesh, 3000)
    return mesh, lon, lat, z


def _same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(
        a[~np.isnan(a)].view(np.int64), b[~np.isnan(b)].view(np.int64))


@pytest.mark.parametrize("shape", ["2d", "mid", "ifc"])
@pytest.mark.parametrize("method", ["nearest", "leastsq", "barycentric", "idw"])
def test_sampling_bitwise(gpu_ctx, world, shape, method):
    mesh, lon, lat, z = world
    K, N = mesh.sigma.size, mesh.n_nodes
    rng = np.random.default_rng(9)
    planes = {"2d": 1, "mid": K - 1, "ifc": K}[shape]
    fa = rng.standard_normal((planes, N)) if planes > 1 else rng.standard_normal(N)
    fb = fa + 0.1 * rng.standard_normal(fa.shape)
    if shape != "ifc":
        fa.reshape(-1)[rng.choice(fa.size, fa.size // 50, replace=False)] = np.nan
    gpu_ctx.upload_mesh(mesh)
    w = np.zeros((K, N), np.float32)
    u = np.zeros((K - 1, N), np.float32)
    for s, f in ((0, fa), (1, fb)):
        gpu_ctx.upload_field(s, u, u, w)
        gpu_ctx.upload_scalar(s, 0, f)
    status = np.where(rng.random(lon.size) < 0.1, 3, 1).astype(np.int32)
    om = OracleMesh(mesh, "kmath")
    nn = om.nearest(lon, lat)
    p = gpu_ctx.particles(lon, lat, z, status=status, nn=nn)
    for alpha in (0.0, 0.3):
        got = gpu_ctx.sample(p, 0, 0, 1, alpha, method)
        want = om.sample(fa, lon, lat, z, status, nn, values_b=fb, alpha=alpha, method=method)
        assert _same(got, want), (shape, method, alpha)
        assert np.isnan(got[status != 1]).all()


@pytest.mark.parametrize("shapes", [("2d", "2d"), ("mid", "mid"), ("2d", "mid"), ("mid", "ifc")])
@pytest.mark.parametrize("method", ["leastsq", "nearest"])
def test_sample_vars_matches_single(gpu_ctx, world, shapes, method):
    """kelp_sample_vars (the one-pass least-squares pair included) gives
    each variable's kelp_sample result bit for bit."""
    mesh, lon, lat, z = world
    K, N = mesh.sigma.size, mesh.n_nodes
    rng = np.random.default_rng(17)
    gpu_ctx.upload_mesh(mesh)
    w = np.zeros((K, N), np.float32)
    u = np.zeros((K - 1, N), np.float32)
    for s in (0, 1):
        gpu_ctx.upload_field(s, u, u, w)
    for var, shape in enumerate(shapes):
        planes = {"2d": 1, "mid": K - 1, "ifc": K}[shape]
        fa = rng.standard_normal((planes, N)) if planes > 1 else rng.standard_normal(N)
        fb = fa + 0.1 * rng.standard_normal(fa.shape)
        if shape != "ifc":
            fa.reshape(-1)[rng.choice(fa.size, fa.size // 20, replace=False)] = np.nan
        gpu_ctx.upload_scalar(0, var, fa)
        gpu_ctx.upload_scalar(1, var, fb)
    status = np.where(rng.random(lon.size) < 0.1, 3, 1).astype(np.int32)
    nn = OracleMesh(mesh, "kmath").nearest(lon, lat)
    p = gpu_ctx.particles(lon, lat, z, status=status, nn=nn)
    for rnd in range(2):
        if rnd == 1:
            # a new record of variable 1 in slot 0: the paired (interleaved)
            # copy of the two columns must follow it
            planes = {"2d": 1, "mid": K - 1, "ifc": K}[shapes[1]]
            fn = rng.standard_normal((planes, N)) if planes > 1 else rng.standard_normal(N)
            gpu_ctx.upload_scalar(0, 1, fn)
        for alpha in (0.0, 0.3):
            both = gpu_ctx.sample_vars(p, (0, 1), 0, 1, alpha, method)
            gpu_ctx.synchronize()
            rows = both.cpu().numpy()
            for var in (0, 1):
                one = gpu_ctx.sample(p, var, 0, 1, alpha, method)
                assert _same(rows[var], one), (shapes, method, alpha, var, rnd)
