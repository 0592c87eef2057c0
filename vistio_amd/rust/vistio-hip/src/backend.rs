//! `HipBackend`: GpuBackend (crates/vistio-gpu/src/backend.rs:19-46) on the
//! device. Each call copies its host buffers in, runs one kernel and copies
//! the result out; length mismatches come back as VistioError::Gpu with the
//! reference's messages, as CpuFallback reports them.

use vistio_gpu::{ComputeBuffer, GpuBackend};
use vistio_types::{VistioError, VistioResult};

use crate::ctx::Ctx;
use crate::ffi::*;

pub struct HipBackend {
    device: i32,
    ctx: Option<Ctx>,
}

impl HipBackend {
    /// A backend for HIP device `device`; `init` creates its context.
    pub fn new(device: i32) -> Self {
        Self { device, ctx: None }
    }

    fn ctx(&self) -> VistioResult<&Ctx> {
        self.ctx.as_ref().ok_or_else(|| VistioError::Gpu("HipBackend used before init()".into()))
    }

    /// The library's status as VistioError::Gpu (the backend's one error kind).
    fn gpu(&self, ctx: &Ctx, rc: VhStatus) -> VistioResult<()> {
        if rc == VH_OK {
            Ok(())
        } else {
            Err(VistioError::Gpu(ctx.last_error()))
        }
    }
}

impl GpuBackend for HipBackend {
    fn init(&mut self) -> VistioResult<()> {
        if self.ctx.is_none() {
            self.ctx = Some(Ctx::new(self.device)?);
        }
        Ok(())
    }

    fn name(&self) -> &str {
        "hip_gfx950"
    }

    /// out[i] = a[i] + scale * b[i] with a separate f32 multiply and add (no
    /// fused multiply-add), as the reference's loop computes it.
    fn axpy(&self, a: &ComputeBuffer, b: &ComputeBuffer, scale: f32, out: &mut ComputeBuffer) -> VistioResult<()> {
        let ctx = self.ctx()?;
        let rc = unsafe {
            vh_backend_axpy(
                ctx.0,
                a.as_slice().as_ptr(),
                a.len() as u64,
                b.as_slice().as_ptr(),
                b.len() as u64,
                scale,
                out.as_mut_slice().as_mut_ptr(),
                out.len() as u64,
            )
        };
        self.gpu(ctx, rc)
    }

    /// Sum of a[i] * b[i] in f64.
    fn dot(&self, a: &ComputeBuffer, b: &ComputeBuffer) -> VistioResult<f64> {
        let ctx = self.ctx()?;
        let mut r = 0.0f64;
        let rc = unsafe {
            vh_backend_dot(ctx.0, a.as_slice().as_ptr(), a.len() as u64, b.as_slice().as_ptr(), b.len() as u64, &mut r)
        };
        self.gpu(ctx, rc)?;
        Ok(r)
    }

    /// The trait's fill cannot report errors; a device failure panics with
    /// the library's message.
    fn fill(&self, buffer: &mut ComputeBuffer, value: f32) {
        let ctx = self.ctx().expect("HipBackend::fill before init()");
        let n = buffer.len() as u64;
        let rc = unsafe { vh_backend_fill(ctx.0, buffer.as_mut_slice().as_mut_ptr(), n, value) };
        if let Err(e) = self.gpu(ctx, rc) {
            panic!("vh_backend_fill: {e}");
        }
    }

    fn is_gpu(&self) -> bool {
        true
    }
}
