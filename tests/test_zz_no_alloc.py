"""kelp_sort_particles_noalloc (the allocation-free re-sort, its scratch in a
static device array) against kelp_sort_particles through the C ABI: two
calls in a row, the second with a smaller input, every permuted array
bit-identical; a particle count above the maximum is refused."""
import ctypes

import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd._native import check

pytestmark = pytest.mark.gpu

MAX_PARTICLES = 1 << 20


def _state(ctx, mesh, n, seed):
    rng = np.random.default_rng(seed)
    lon, lat, z = synth.wet_area_seeds(mesh, n, seed=seed)
    status = rng.choice([1, 1, 1, 2, 3, 4], n).astype(np.int32)
    pid = rng.permutation(n).astype(np.int64) + 17
    p = ctx.particles(lon, lat, z, status=status, pid=pid, ws=rng.integers(0, 5, n))
    t = ctx.torch
    p.u.copy_(t.as_tensor(rng.normal(size=n), device=p.u.device))
    p.fr.copy_(t.as_tensor(rng.integers(0, 3, n).astype(np.int8), device=p.fr.device))
    return p


def _host(p):
    return {k: getattr(p, k).cpu().numpy() for k in
            ("lon", "lat", "z", "status", "nn", "fr", "u", "v", "w", "pid", "ws")}


def test_noalloc_sort_matches_allocating_sort(gpu_ctx):
    mesh = synth.sphere_mesh(20000)
    gpu_ctx.upload_mesh(mesh)
    lib = gpu_ctx.lib
    for n, seed in ((200_000, 1), (7_777, 2)):
        a = _state(gpu_ctx, mesh, n, seed)
        b = _state(gpu_ctx, mesh, n, seed)
        gpu_ctx.torch.cuda.current_stream().synchronize()
        sa, sb = a.as_struct(), b.as_struct()
        check(lib.kelp_sort_particles(gpu_ctx.h, ctypes.byref(sa)), "kelp_sort_particles")
        check(lib.kelp_sort_particles_noalloc(gpu_ctx.h, ctypes.byref(sb)), "kelp_sort_particles_noalloc")
        gpu_ctx.synchronize()
        ha, hb = _host(a), _host(b)
        assert not np.array_equal(ha["pid"], np.sort(ha["pid"])), "the sort should move particles"
        for k in ha:
            assert np.array_equal(ha[k].view(np.uint8), hb[k].view(np.uint8)), (n, k)


def test_noalloc_sort_refuses_above_maximum(gpu_ctx):
    mesh = synth.sphere_mesh(2000)
    gpu_ctx.upload_mesh(mesh)
    p = _state(gpu_ctx, mesh, 64, 3)
    s = p.as_struct()
    s.n = MAX_PARTICLES + 1
    rc = gpu_ctx.lib.kelp_sort_particles_noalloc(gpu_ctx.h, ctypes.byref(s))
    assert rc != 0
    assert b"exceed" in gpu_ctx.lib.kelp_last_error()
