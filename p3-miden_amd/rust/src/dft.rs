//! `TwoAdicSubgroupDft` on the GPU (the LDE called at
//! p3-miden-lifted-stark/src/prover/commit.rs:217-221).

use std::sync::Arc;

use p3_dft::{Radix2DitParallel, TwoAdicSubgroupDft};
use p3_field::PrimeField64;
use p3_goldilocks::Goldilocks;
use p3_matrix::Matrix;
use p3_matrix::dense::RowMajorMatrix;

use crate::context::GpuContext;
use crate::ffi;

/// Coset LDEs on the device; the plain DFT (`dft_batch`, the trait's one required
/// method, not on the prover's hot path) stays on the CPU.  Values are unique, so
/// the result equals Radix2DitParallel's bit for bit.
#[derive(Clone)]
pub struct GpuDft {
    ctx: Arc<GpuContext>,
    cpu: Radix2DitParallel<Goldilocks>,
}

impl GpuDft {
    pub fn new(ctx: Arc<GpuContext>) -> Self {
        Self {
            ctx,
            cpu: Radix2DitParallel::default(),
        }
    }
}

impl Default for GpuDft {
    fn default() -> Self {
        Self::new(GpuContext::default_device())
    }
}

pub(crate) fn to_words(values: &[Goldilocks]) -> Vec<u64> {
    // the field's internal form may be non-canonical; the library takes canonical words
    values.iter().map(|x| x.as_canonical_u64()).collect()
}

pub(crate) fn from_words(words: &[u64]) -> Vec<Goldilocks> {
    words.iter().map(|&x| Goldilocks::new(x)).collect()
}

impl TwoAdicSubgroupDft<Goldilocks> for GpuDft {
    type Evaluations = RowMajorMatrix<Goldilocks>;

    fn dft_batch(&self, mat: RowMajorMatrix<Goldilocks>) -> Self::Evaluations {
        self.cpu.dft_batch(mat).to_row_major_matrix()
    }

    /// Evaluations of each column's polynomial on shift * K (|K| = height << added_bits),
    /// in natural order, as the trait promises.  `commit_traces` bit-reverses them
    /// afterwards; `crate::commit_traces` does both on the device in one call.
    fn coset_lde_batch(
        &self,
        mat: RowMajorMatrix<Goldilocks>,
        added_bits: usize,
        shift: Goldilocks,
    ) -> Self::Evaluations {
        let (n, w) = (mat.height(), mat.width());
        let words = to_words(&mat.values);
        let mut out = vec![0u64; (n << added_bits) * w];
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: `words` holds n * w words and `out` (n << added_bits) * w, as the call requires.
        let rc = unsafe {
            ffi::p3g_coset_lde_batch(
                self.ctx.raw(),
                words.as_ptr(),
                n as u64,
                w as u64,
                added_bits as u32,
                shift.as_canonical_u64(),
                0,
                out.as_mut_ptr(),
            )
        };
        ffi::check(rc, "p3g_coset_lde_batch");
        RowMajorMatrix::new(from_words(&out), w)
    }
}
