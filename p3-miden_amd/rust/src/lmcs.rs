//! `Lmcs` / `LmcsTree` (p3-miden-lmcs/src/lib.rs:176-295) with the tree built and
//! opened on the GPU.  Hashing single rows, compression and the verifier side
//! (`open_batch`, `read_batch_proof_from_channel`) stay with the CPU configuration
//! the GPU one wraps: the hint stream the GPU writes is the one it reads.

use std::sync::{Arc, OnceLock};

use p3_field::PrimeField64;
use p3_goldilocks::Goldilocks;
use p3_matrix::Matrix;
use p3_miden_lmcs::{Lmcs, LmcsError, LmcsTree, OpenedRows, RowList};
use p3_miden_transcript::{ProverChannel, VerifierChannel};

use crate::channel::{Digest, bridge};
use crate::context::GpuContext;
use crate::dft::from_words;
use crate::ffi;

/// The GPU form of a Goldilocks Poseidon2-GL12 `LmcsConfig` (StatefulSponge<Poseidon2Goldilocks<12>,
/// 12, 8, 4> + TruncatedPermutation, p3-miden-dev-utils/src/configs.rs:194-218).  `inner` is that
/// CPU configuration; the permutation must be the one the library's kernels compute
/// (new_from_rng_128 with SmallRng seed 2025).
#[derive(Clone)]
pub struct GpuLmcs<L> {
    pub inner: L,
    pub(crate) ctx: Arc<GpuContext>,
}

impl<L> GpuLmcs<L> {
    pub fn new(inner: L, ctx: Arc<GpuContext>) -> Self {
        Self { inner, ctx }
    }
}

/// A tree whose leaves and digest layers live in HBM.  It keeps the host matrices it
/// was built from, because `LmcsTree::leaves` hands out references to them.
pub struct GpuTree<M> {
    pub(crate) raw: *mut ffi::p3g_tree,
    pub(crate) ctx: Arc<GpuContext>,
    // host copies of the committed matrices: given at build time, or (for trees whose
    // matrices were made on the device, such as the LDEs of commit_traces) downloaded
    // on the first call to `leaves`
    leaves: OnceLock<Vec<M>>,
    loader: Option<Box<dyn Fn(&Self) -> Vec<M> + Send + Sync>>,
    root: [Goldilocks; 4],
    height: usize,
    alignment: usize,
    widths: Vec<usize>,
}

// SAFETY: the tree handle is only used under the context's lock.
unsafe impl<M: Send> Send for GpuTree<M> {}
unsafe impl<M: Sync> Sync for GpuTree<M> {}

impl<M> Drop for GpuTree<M> {
    fn drop(&mut self) {
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: the handle came from p3g_lmcs_build or p3g_commit_traces and is freed once.
        unsafe { ffi::p3g_tree_free(self.raw) }
    }
}

impl<M> GpuTree<M> {
    pub(crate) fn from_raw(raw: *mut ffi::p3g_tree, ctx: Arc<GpuContext>, leaves: Vec<M>) -> Self {
        let t = Self::bare(raw, ctx);
        let _ = t.leaves.set(leaves);
        t
    }

    pub(crate) fn with_loader(
        raw: *mut ffi::p3g_tree,
        ctx: Arc<GpuContext>,
        loader: Box<dyn Fn(&Self) -> Vec<M> + Send + Sync>,
    ) -> Self {
        let mut t = Self::bare(raw, ctx);
        t.loader = Some(loader);
        t
    }

    fn bare(raw: *mut ffi::p3g_tree, ctx: Arc<GpuContext>) -> Self {
        let mut root = [0u64; 4];
        // SAFETY: `raw` is a live tree; `root` and `w` have the sizes the calls write.
        let (height, alignment, widths) = unsafe {
            ffi::check(ffi::p3g_tree_root(raw, root.as_mut_ptr()), "p3g_tree_root");
            let k = ffi::p3g_tree_num_matrices(raw) as usize;
            let mut w = vec![0u64; k];
            ffi::check(ffi::p3g_tree_widths(raw, w.as_mut_ptr()), "p3g_tree_widths");
            (
                ffi::p3g_tree_height(raw) as usize,
                ffi::p3g_tree_alignment(raw) as usize,
                w.into_iter().map(|x| x as usize).collect(),
            )
        };
        Self {
            raw,
            ctx,
            leaves: OnceLock::new(),
            loader: None,
            root: core::array::from_fn(|i| Goldilocks::new(root[i])),
            height,
            alignment,
            widths,
        }
    }
}

impl<L> GpuLmcs<L> {
    fn build<M: Matrix<Goldilocks>>(&self, leaves: Vec<M>, alignment: u32) -> GpuTree<M> {
        // row-major canonical words of every matrix, in the order given (ascending heights)
        let words: Vec<Vec<u64>> = leaves
            .iter()
            .map(|m| m.rows().flat_map(|row| row.into_iter().map(|x| x.as_canonical_u64())).collect())
            .collect();
        let mats: Vec<ffi::p3g_matrix> = leaves
            .iter()
            .zip(&words)
            .map(|(m, w)| ffi::p3g_matrix {
                data: w.as_ptr(),
                height: m.height() as u64,
                width: m.width() as u64,
            })
            .collect();
        let mut raw = core::ptr::null_mut();
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: every descriptor points at height * width words that outlive the call
        // (host matrices are copied to the device before it returns).
        let rc = unsafe {
            ffi::p3g_lmcs_build(
                self.ctx.raw(),
                ffi::P3G_HASH_POSEIDON2_GL12,
                mats.as_ptr(),
                mats.len() as u64,
                alignment,
                0,
                &mut raw,
            )
        };
        ffi::check(rc, "p3g_lmcs_build");
        drop(_guard);
        GpuTree::from_raw(raw, self.ctx.clone(), leaves)
    }
}

impl<L> Lmcs for GpuLmcs<L>
where
    L: Lmcs<F = Goldilocks, Commitment = Digest>,
{
    type F = Goldilocks;
    type Commitment = Digest;
    type BatchProof = L::BatchProof;
    type Tree<M: Matrix<Goldilocks>> = GpuTree<M>;

    fn build_tree<M: Matrix<Goldilocks>>(&self, leaves: Vec<M>) -> GpuTree<M> {
        self.build(leaves, 1)
    }

    fn build_aligned_tree<M: Matrix<Goldilocks>>(&self, leaves: Vec<M>) -> GpuTree<M> {
        // SAFETY: a pure function of its argument
        let a = unsafe { ffi::p3g_lmcs_alignment(ffi::P3G_HASH_POSEIDON2_GL12) };
        debug_assert_eq!(a as usize, self.inner.alignment());
        self.build(leaves, a)
    }

    fn hash<'a, I>(&self, rows: I) -> Digest
    where
        I: IntoIterator<Item = &'a [Goldilocks]>,
        Goldilocks: 'a,
    {
        self.inner.hash(rows)
    }

    fn compress(&self, left: Digest, right: Digest) -> Digest {
        self.inner.compress(left, right)
    }

    fn open_batch<Ch>(
        &self,
        commitment: &Digest,
        widths: &[usize],
        log_max_height: u8,
        indices: impl IntoIterator<Item = usize>,
        channel: &mut Ch,
    ) -> Result<OpenedRows<Goldilocks>, LmcsError>
    where
        Ch: VerifierChannel<F = Goldilocks, Commitment = Digest>,
    {
        self.inner.open_batch(commitment, widths, log_max_height, indices, channel)
    }

    fn read_batch_proof_from_channel<Ch>(
        &self,
        widths: &[usize],
        log_max_height: u8,
        indices: &[usize],
        channel: &mut Ch,
    ) -> Result<Self::BatchProof, LmcsError>
    where
        Ch: VerifierChannel<F = Goldilocks, Commitment = Digest>,
    {
        self.inner.read_batch_proof_from_channel(widths, log_max_height, indices, channel)
    }

    fn alignment(&self) -> usize {
        self.inner.alignment()
    }
}

impl<M: Matrix<Goldilocks>> LmcsTree<Goldilocks, Digest, M> for GpuTree<M> {
    fn root(&self) -> Digest {
        Digest::from(self.root)
    }

    fn height(&self) -> usize {
        self.height
    }

    fn leaves(&self) -> &[M] {
        self.leaves.get_or_init(|| match &self.loader {
            Some(load) => load(self),
            None => Vec::new(),
        })
    }

    fn rows(&self, index: usize) -> RowList<Goldilocks> {
        let total: usize = self.widths.iter().sum();
        let mut out = vec![0u64; total];
        let mut len = 0u64;
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: `out` holds the sum of the aligned widths, which is what the call writes.
        let rc = unsafe { ffi::p3g_tree_rows(self.raw, index as u64, out.as_mut_ptr(), total as u64, &mut len) };
        ffi::check(rc, "p3g_tree_rows");
        RowList::new(from_words(&out[..len as usize]), &self.widths)
    }

    fn alignment(&self) -> usize {
        self.alignment
    }

    fn widths(&self) -> Vec<usize> {
        self.widths.clone()
    }

    fn prove_batch<Ch>(&self, indices: impl IntoIterator<Item = usize>, channel: &mut Ch)
    where
        Ch: ProverChannel<F = Goldilocks, Commitment = Digest>,
    {
        let idx: Vec<u64> = indices.into_iter().map(|i| i as u64).collect();
        let cb = bridge(channel);
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: `cb` borrows `channel` for the duration of the call only.
        let rc = unsafe { ffi::p3g_tree_prove_batch_cb(self.raw, idx.as_ptr(), idx.len() as u64, &cb) };
        ffi::check(rc, "p3g_tree_prove_batch_cb");
    }
}
