"""nb_copy_vertices_noalloc, the entry point for callers that cannot allocate
device memory, against nb_copy_vertices on the same state: bit-identical
float4 vertices (positions and velocities, each alone and both), on two calls
in a row where the second has fewer bodies (the staging array keeps the first
call's contents), the first spanning several 65,536-body chunks with a
partial last one."""
import numpy as np
import pytest

from helpers import make_sim

pytestmark = pytest.mark.gpu


def _nan4(n):
    return np.full((n, 4), np.nan, np.float32)


def test_copy_vertices_noalloc_matches_copy_vertices(nb):
    from nbody_amd import universe as U

    for n, seed in ((300_000, 61), (5_000, 62)):
        st = U.plummer(n, seed=seed)
        sim = make_sim(st, theta=0.5, softening=0.01, dt=1e-3, vote_width=64,
                       flags=nb._lib.NB_FP64_POSITIONS | nb._lib.NB_FUSED_UPDATE)
        sim.steps(2)
        ref_p, ref_v = _nan4(n), _nan4(n)
        sim.copy_vertices(ref_p, ref_v)
        assert not np.isnan(ref_p).any() and (ref_p[:, 3] == 1.0).all() and (ref_v[:, 3] == 1.0).all()
        p, v = _nan4(n), _nan4(n)
        sim.copy_vertices_noalloc(p, v)
        assert np.array_equal(p.view(np.uint32), ref_p.view(np.uint32)), n
        assert np.array_equal(v.view(np.uint32), ref_v.view(np.uint32)), n
        p = _nan4(n)
        sim.copy_vertices_noalloc(p, None)
        assert np.array_equal(p.view(np.uint32), ref_p.view(np.uint32)), n
        v = _nan4(n)
        sim.copy_vertices_noalloc(None, v)
        assert np.array_equal(v.view(np.uint32), ref_v.view(np.uint32)), n
        sim.copy_vertices_noalloc(None, None)
        sim.close()
