"""NarrowPhase on the GPU (vh_narrow_vertex_triangle, HipVertexTriangleTest)
against the oracle's VertexTriangleTest (vertex_triangle.rs:25-152): both
run the reference's f32 arithmetic in its operation order without FMA, so
every contact must match bit for bit, in the reference's order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(c, seed, fold=True):
    from oracle import mesh
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    rng = np.random.default_rng(seed)
    pz = (rng.normal(size=px.size) * 0.004).astype(np.float32)
    if fold:  # fold the right half back over the left, 6 mm above it
        right = px > 0.0
        px = px.copy()
        px[right] = -px[right]
        pz[right] += 0.006
    return np.stack([px, py, pz]).astype(np.float32), idx


def _near_pairs(pos, r):
    d2 = ((pos[:, :, None] - pos[:, None, :]) ** 2).sum(0)
    a, b = np.nonzero(d2 < r * r)
    keep = a != b
    return np.stack([a[keep], b[keep]], 1).astype(np.uint32)


def _fields(c):
    return c.view(np.uint32).reshape(len(c), -1) if len(c) else np.zeros((0, 14), np.uint32)


@pytest.mark.parametrize("c,seed,thick", [(16, 0, 0.01), (24, 1, 0.008), (12, 2, 0.02)])
def test_matches_oracle_bitwise(c, seed, thick):
    import vistio_amd as va
    from oracle import ref
    pos, idx = _scene(c, seed)
    cand = _near_pairs(pos, 0.05)
    want = ref.narrow_vt(pos[0], pos[1], pos[2], idx, cand, thick)
    nt = va.HipVertexTriangleTest()
    got = nt.detect(cand, pos, idx, thick)
    print(f"{c}x{c}: {len(cand)} candidates, {len(want)} contacts", flush=True)
    assert len(want) > 50
    assert len(got) == len(want)
    g, w = _fields(got), _fields(want)
    assert np.array_equal(g[:, :11], w[:, :11])
    assert (got["contact_type"] == 0).all() and (got["is_self"] == 1).all()
    assert nt.name() == "vertex_triangle_test"  # contact_tests.rs:366-369
    nt.close()


def test_cap_empty_and_errors():
    import ctypes as C

    import vistio_amd as va
    from vistio_amd import _native as nat
    from oracle import ref
    pos, idx = _scene(10, 3)
    cand = _near_pairs(pos, 0.06)
    want = ref.narrow_vt(pos[0], pos[1], pos[2], idx, cand, 0.02)
    nt = va.HipVertexTriangleTest()
    L = nat.load()
    ca, cb = np.ascontiguousarray(cand[:, 0]), np.ascontiguousarray(cand[:, 1])
    px, py, pz = (np.ascontiguousarray(pos[k]) for k in range(3))
    out = np.zeros(7, dtype=va.narrow.CONTACT_DTYPE)
    cnt = C.c_uint64(0)
    st = L.vh_narrow_vertex_triangle(nt._ctx, px.size, nat.ptr(px, C.c_float), nat.ptr(py, C.c_float),
                                     nat.ptr(pz, C.c_float), idx.size // 3, nat.ptr(idx, C.c_uint32), ca.size,
                                     nat.ptr(ca, C.c_uint32), nat.ptr(cb, C.c_uint32), 0.02,
                                     out.ctypes.data_as(C.POINTER(nat.ContactPairC)), 7, C.byref(cnt))
    assert st == 0 and cnt.value == len(want) > 7
    assert np.array_equal(_fields(out)[:, :11], _fields(want[:7])[:, :11])
    assert len(nt.detect(np.zeros((0, 2), np.uint32), pos, idx, 0.02)) == 0
    bad = cand[:3].copy()
    bad[1, 0] = pos.shape[1]
    with pytest.raises(va.VistioError):
        nt.detect(bad, pos, idx, 0.02)
    bad_idx = idx.copy()
    bad_idx[4] = pos.shape[1] + 5
    with pytest.raises(va.InvalidMesh):
        nt.detect(cand[:3], pos, bad_idx, 0.02)
    # the same context serves calls of other sizes (its buffers grow)
    pos2, idx2 = _scene(20, 4)
    cand2 = _near_pairs(pos2, 0.04)
    want2 = ref.narrow_vt(pos2[0], pos2[1], pos2[2], idx2, cand2, 0.01)
    got2 = nt.detect(cand2, pos2, idx2, 0.01)
    assert np.array_equal(_fields(got2)[:, :11], _fields(want2)[:, :11])
    nt.close()
