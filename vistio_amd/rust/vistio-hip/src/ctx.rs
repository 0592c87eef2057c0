//! Ownership of one `vh_ctx` and the status -> VistioError mapping.

use std::ffi::CStr;

use vistio_types::{VistioError, VistioResult};

use crate::ffi::*;

/// One `vh_ctx`: device memory and one HIP stream, destroyed on drop.
pub(crate) struct Ctx(pub(crate) *mut VhCtx);

// A context is used by one thread at a time: its owners take &mut self for
// every call that changes device state, and none of them is Sync.
unsafe impl Send for Ctx {}

impl Ctx {
    pub(crate) fn new(device: i32) -> VistioResult<Self> {
        let mut p = std::ptr::null_mut();
        let rc = unsafe { vh_create(device, &mut p) };
        if rc != VH_OK || p.is_null() {
            return Err(VistioError::Gpu(format!("vh_create(device {device}) failed with status {rc}")));
        }
        Ok(Ctx(p))
    }

    pub(crate) fn last_error(&self) -> String {
        unsafe {
            let m = vh_last_error(self.0);
            if m.is_null() {
                String::new()
            } else {
                CStr::from_ptr(m).to_string_lossy().into_owned()
            }
        }
    }

    /// VH_E_INVALID_MESH -> InvalidMesh; VH_E_GPU and VH_E_INTERNAL -> Gpu;
    /// every other failure -> InvalidConfig (not initialised, factorization
    /// failure, unsupported model, bad argument, non-finite state), with the
    /// library's message.
    pub(crate) fn check(&self, rc: VhStatus) -> VistioResult<()> {
        if rc == VH_OK {
            return Ok(());
        }
        let msg = self.last_error();
        Err(match rc {
            VH_E_INVALID_MESH => VistioError::InvalidMesh(msg),
            VH_E_GPU | VH_E_INTERNAL => VistioError::Gpu(msg),
            _ => VistioError::InvalidConfig(msg),
        })
    }
}

impl Drop for Ctx {
    fn drop(&mut self) {
        if !self.0.is_null() {
            unsafe {
                vh_destroy(self.0);
            }
        }
    }
}
