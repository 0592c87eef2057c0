"""GpuBackend on the MI355X: the reference's own backend tests
(crates/vistio-gpu/tests/gpu_tests.rs) run against HipBackend, plus large
seeded vectors against numpy (axpy bit-exact, dot to f64 rounding of the
partial sums). ComputeBuffer's own tests run on the CPU."""
import numpy as np
import pytest


def test_buffer_kats():
    from vistio_amd import ComputeBuffer
    b = ComputeBuffer.zeros(100)                                     # gpu_tests.rs:8-13
    assert len(b) == 100 and np.all(b.as_slice() == 0.0)
    b = ComputeBuffer.from_data([1.0, 2.0, 3.0])                     # :15-20
    assert len(b) == 3 and b.as_slice().tolist() == [1.0, 2.0, 3.0]
    b = ComputeBuffer.zeros(3)                                       # :22-27
    b.copy_from_slice([7.0, 8.0, 9.0])
    assert b.as_slice().tolist() == [7.0, 8.0, 9.0]


@pytest.mark.gpu
def test_hip_backend_reference_kats():
    import vistio_amd as va
    from vistio_amd import ComputeBuffer
    be = va.HipBackend()
    be.init()                                                        # cpu_init, :31-37
    assert be.name() == "hip_gfx950" and be.is_gpu()
    a, b = ComputeBuffer.from_data([1.0, 2.0, 3.0]), ComputeBuffer.from_data([4.0, 5.0, 6.0])
    out = ComputeBuffer.zeros(3)
    be.axpy(a, b, 2.0, out)                                          # cpu_axpy, :39-48
    assert out.as_slice().tolist() == [9.0, 12.0, 15.0]
    with pytest.raises(va.GpuError, match="AXPY buffer length mismatch"):  # :50-57
        be.axpy(ComputeBuffer.from_data([1.0, 2.0]), b, 1.0, out)
    assert abs(be.dot(a, b) - 32.0) < 1e-10                          # cpu_dot, :59-66
    buf = ComputeBuffer.zeros(5)
    be.fill(buf, np.pi)                                              # cpu_fill, :68-73
    assert np.all(np.abs(buf.as_slice() - np.float32(np.pi)) < 1e-6)
    be.close()


@pytest.mark.gpu
def test_hip_backend_large_vectors():
    import vistio_amd as va
    from vistio_amd import ComputeBuffer
    be = va.HipBackend()
    be.init()
    rng = np.random.default_rng(3)
    for n in (1, 255, 4097, 3_000_001):
        a = ComputeBuffer(rng.standard_normal(n).astype(np.float32))
        b = ComputeBuffer(rng.standard_normal(n).astype(np.float32))
        s = np.float32(rng.standard_normal())
        out = ComputeBuffer.zeros(n)
        be.axpy(a, b, float(s), out)
        want = a.data + s * b.data            # f32 multiply then f32 add, as the reference
        assert np.array_equal(out.data, want)
        exact = float(np.sum(a.data.astype(np.float64) * b.data.astype(np.float64)))
        scale = float(np.sum(np.abs(a.data.astype(np.float64) * b.data.astype(np.float64))))
        assert abs(be.dot(a, b) - exact) <= 1e-14 * max(scale, 1.0)
        be.fill(out, -2.5)
        assert np.all(out.data == np.float32(-2.5))
    be.close()
