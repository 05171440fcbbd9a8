"""zdx_contact_pairs_noalloc against zdx_contact_pairs through the C ABI:
bitwise the same per-pair results, on two calls in a row (the second with a
smaller batch, so stale contents of the shared scratch array from the first
call would show), with sensitivities on the second; a call too large for the
scratch array fails with ZDX_CAPACITY and leaves the context usable; a mesh
registered after the last allocating call is refused (ZDX_BAD_ARG)."""
import numpy as np
import pytest

from zdx import workloads as W

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu

KEYS = ("tri_checks", "n_segments", "n_comps", "n_loops_raw", "loop_offset", "seg_offset", "loops", "vtx_offset",
        "loop_vtx", "segments")


def _args(ids, pairs):
    return ([ids[p[0]] for p in pairs], [ids[p[1]] for p in pairs], [p[2] for p in pairs], [p[3] for p in pairs],
            [p[4] for p in pairs], [p[5] for p in pairs])


def test_noalloc_matches_contact_pairs_twice():
    ctx = zdx.Context(0)
    meshes, names, pairs = W.uneven_pairs(256, seed=9)
    ids = [ctx.register_mesh(m) for m in meshes]
    first = _args(ids, pairs)
    second = _args(ids, pairs[32:96])
    ref1 = ctx.contact_pairs(*first, split=True, mode=zdx.MODE_BATCHED, k=1.5, segments=True)
    ref2 = ctx.contact_pairs(*second, split=False, mode=zdx.MODE_PLAIN, k=0.5, segments=True, sensitivities=True)
    got1 = ctx.contact_pairs(*first, split=True, mode=zdx.MODE_BATCHED, k=1.5, segments=True, noalloc=True)
    print(f"[noalloc] first call: {len(got1['loops'])} loops, {len(got1['segments'])} segments", flush=True)
    got2 = ctx.contact_pairs(*second, split=False, mode=zdx.MODE_PLAIN, k=0.5, segments=True, sensitivities=True,
                             noalloc=True)
    print(f"[noalloc] second call: {len(got2['loops'])} loops", flush=True)
    for ref, got in ((ref1, got1), (ref2, got2)):
        for key in KEYS:
            assert np.array_equal(ref[key], got[key]), key
    assert np.array_equal(ref2["loop_jac"], got2["loop_jac"])
    assert len(got1["loops"]) > 100 and len(got2["loops"]) > 20
    ctx.close()


def test_noalloc_refuses_what_it_cannot_hold():
    ctx = zdx.Context(0)
    meshes, names, pairs = W.uneven_pairs(4096, seed=2)
    ids = [ctx.register_mesh(m) for m in meshes]
    ctx.contact_pairs(*_args(ids, pairs[:4]))          # uploads the meshes
    with pytest.raises(zdx.ZdxError) as e:
        ctx.contact_pairs(*_args(ids, pairs + pairs), noalloc=True)
    assert e.value.status == zdx.CAPACITY
    small = _args(ids, pairs[:16])
    assert np.array_equal(ctx.contact_pairs(*small, noalloc=True)["loops"], ctx.contact_pairs(*small)["loops"])
    late = ctx.register_mesh(meshes[0] * 1.5)
    with pytest.raises(zdx.ZdxError) as e:
        ctx.contact_pairs([late], [ids[0]], [[0, 0, 0]], [[1, 0, 0, 0]], [[2.0, 0, 0]], [[1, 0, 0, 0]], noalloc=True)
    assert e.value.status == zdx.BAD_ARG
    ctx.close()
