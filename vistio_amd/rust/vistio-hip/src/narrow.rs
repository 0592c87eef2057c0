//! `NarrowPhase` (narrow.rs:21-35) on the GPU: `VertexTriangleTest`
//! (vertex_triangle.rs:17-82) through `vh_narrow_vertex_triangle`, in the
//! reference's f32 arithmetic and contact order.

use std::sync::Mutex;

use vistio_contact::{
    broad::CandidatePair,
    contact::{ContactPair, ContactType},
    narrow::NarrowPhase,
};
use vistio_mesh::TriangleMesh;
use vistio_solver::state::SimulationState;
use vistio_types::{VistioError, VistioResult};

use crate::ctx::Ctx;
use crate::ffi::*;

/// The HIP `VertexTriangleTest`. `detect` takes `&self`, so the device
/// context sits behind a mutex (one call at a time per instance).
pub struct HipNarrowPhase {
    ctx: Mutex<Ctx>,
}

impl HipNarrowPhase {
    pub fn new(device: i32) -> VistioResult<Self> {
        Ok(Self { ctx: Mutex::new(Ctx::new(device)?) })
    }
}

impl NarrowPhase for HipNarrowPhase {
    fn detect(
        &self,
        candidates: &[CandidatePair],
        state: &SimulationState,
        mesh: &TriangleMesh,
        thickness: f32,
    ) -> VistioResult<Vec<ContactPair>> {
        let ctx = self.ctx.lock().map_err(|_| VistioError::Gpu("narrow phase context poisoned".into()))?;
        let n = state.pos_x.len();
        if state.pos_y.len() != n || state.pos_z.len() != n {
            return Err(VistioError::InvalidMesh(format!("positions must have {n} entries")));
        }
        let ca: Vec<u32> = candidates.iter().map(|c| c.a).collect();
        let cb: Vec<u32> = candidates.iter().map(|c| c.b).collect();
        let mut out = vec![VhContactPair::default(); 4 * candidates.len().max(4)];
        let mut count = 0u64;
        let mut call = |out: &mut Vec<VhContactPair>, count: &mut u64| {
            ctx.check(unsafe {
                vh_narrow_vertex_triangle(
                    ctx.0,
                    n as u32,
                    state.pos_x.as_ptr(),
                    state.pos_y.as_ptr(),
                    state.pos_z.as_ptr(),
                    mesh.triangle_count() as u32,
                    mesh.indices.as_ptr(),
                    ca.len() as u64,
                    ca.as_ptr(),
                    cb.as_ptr(),
                    thickness,
                    out.as_mut_ptr(),
                    out.len() as u64,
                    count,
                )
            })
        };
        call(&mut out, &mut count)?;
        if count as usize > out.len() {
            // more contacts than the first guess: again, with room for all
            out.resize(count as usize, VhContactPair::default());
            call(&mut out, &mut count)?;
        }
        out.truncate(count as usize);
        Ok(out
            .iter()
            .map(|c| ContactPair {
                contact_type: ContactType::VertexTriangle,
                indices: c.indices,
                distance: c.distance,
                normal: c.normal,
                barycentric: c.barycentric,
                is_self: c.is_self != 0,
            })
            .collect())
    }

    fn name(&self) -> &str {
        "vertex_triangle_test"
    }
}
