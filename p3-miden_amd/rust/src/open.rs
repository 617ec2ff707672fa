//! The PCS opening (p3-miden-lifted-fri/src/prover.rs:35-116) and the trace commitment
//! (p3-miden-lifted-stark/src/prover/commit.rs:175-228) through the library.  The
//! reference's `open_with_channel` is a free generic function, not a trait method
//! (`deep/mod.rs:54`, `fri/mod.rs:14` keep DEEP and FRI crate-private), so a prover
//! that wants the GPU opening calls this one, with the same arguments, where
//! `prove_multi` calls the reference's (`prover/mod.rs:441-449`).

use p3_field::{BasedVectorSpace, PrimeField64};
use p3_field::extension::BinomialExtensionField;
use p3_goldilocks::Goldilocks;
use p3_matrix::Matrix;
use p3_matrix::dense::RowMajorMatrix;
use p3_miden_lifted_fri::{PcsParams, PcsParamsError};
use p3_miden_lmcs::{Lmcs, LmcsTree};

use crate::channel::{Digest, GpuChannel, bridge};
use crate::dft::{from_words, to_words};
use crate::ffi;
use crate::lmcs::{GpuLmcs, GpuTree};

pub type Challenge = BinomialExtensionField<Goldilocks, 2>;

/// `PcsParams` together with the seven numbers it was made from: the reference keeps
/// the grinding bits and the FRI arity crate-private, and the library needs them.
#[derive(Clone, Debug)]
pub struct GpuPcsParams {
    pub params: PcsParams,
    pub(crate) raw: ffi::p3g_pcs_params,
}

impl GpuPcsParams {
    /// Same arguments and checks as `PcsParams::new` (params.rs:54-88).
    pub fn new(
        log_blowup: u8,
        log_folding_arity: u8,
        log_final_degree: u8,
        folding_pow_bits: usize,
        deep_pow_bits: usize,
        num_queries: usize,
        query_pow_bits: usize,
    ) -> Result<Self, PcsParamsError> {
        let params = PcsParams::new(
            log_blowup,
            log_folding_arity,
            log_final_degree,
            folding_pow_bits,
            deep_pow_bits,
            num_queries,
            query_pow_bits,
        )?;
        Ok(Self {
            params,
            raw: ffi::p3g_pcs_params {
                log_blowup: log_blowup as u32,
                log_folding_arity: log_folding_arity as u32,
                log_final_degree: log_final_degree as u32,
                folding_pow_bits: folding_pow_bits as u32,
                deep_pow_bits: deep_pow_bits as u32,
                num_queries: num_queries as u32,
                query_pow_bits: query_pow_bits as u32,
            },
        })
    }
}

/// `open_with_channel` with DEEP, FRI and the query openings on the GPU and every
/// transcript operation on `channel`, in the reference's order.  Panics where the
/// reference panics (no trees, mixed LDE heights or alignments, N outside 1..=4).
pub fn open_with_channel<L, M, Ch, const N: usize>(
    params: &GpuPcsParams,
    lmcs: &GpuLmcs<L>,
    log_lde_height: u8,
    eval_points: [Challenge; N],
    trace_trees: &[&GpuTree<M>],
    channel: &mut Ch,
) where
    L: Lmcs<F = Goldilocks, Commitment = Digest>,
    M: Matrix<Goldilocks>,
    Ch: GpuChannel,
{
    const { assert!(N > 0, "at least one evaluation point required") };
    assert!(!trace_trees.is_empty(), "at least one trace tree required");
    let z: Vec<u64> = eval_points
        .iter()
        .flat_map(|e| e.as_basis_coefficients_slice().iter().map(|c| c.as_canonical_u64()))
        .collect();
    let trees: Vec<*mut ffi::p3g_tree> = trace_trees.iter().map(|t| t.raw).collect();
    let cb = bridge(channel);
    let _guard = lmcs.ctx.lock.lock().unwrap();
    // SAFETY: z holds 2N words, trees ntrees live handles of this context, and cb borrows
    // the channel for the duration of the call only.
    let rc = unsafe {
        ffi::p3g_open_with_callbacks(
            lmcs.ctx.raw(),
            &params.raw,
            log_lde_height as u32,
            z.as_ptr(),
            N as u32,
            trees.as_ptr(),
            trees.len() as u64,
            &cb,
        )
    };
    ffi::check(rc, "p3g_open_with_callbacks");
}

/// `commit_traces` (commit.rs:175-228) in one call: each trace's coset LDE on the nested
/// shift GENERATOR^(2^(log n_max - log n_j)), bit-reversed rows, and one aligned tree over
/// all LDEs, which stay in HBM.  `leaves()` downloads the bit-reversed LDEs on first use
/// (the GPU opening never needs them on the host).  Traces sorted by ascending height.
pub fn commit_traces<L>(
    lmcs: &GpuLmcs<L>,
    traces: &[RowMajorMatrix<Goldilocks>],
    log_blowup: u8,
) -> GpuTree<RowMajorMatrix<Goldilocks>>
where
    L: Lmcs<F = Goldilocks, Commitment = Digest>,
{
    let words: Vec<Vec<u64>> = traces.iter().map(|t| to_words(&t.values)).collect();
    let mats: Vec<ffi::p3g_matrix> = traces
        .iter()
        .zip(&words)
        .map(|(t, w)| ffi::p3g_matrix {
            data: w.as_ptr(),
            height: t.height() as u64,
            width: t.width() as u64,
        })
        .collect();
    let mut raw = core::ptr::null_mut();
    let guard = lmcs.ctx.lock.lock().unwrap();
    // SAFETY: every descriptor points at height * width words that outlive the call.
    let rc = unsafe {
        ffi::p3g_commit_traces(
            lmcs.ctx.raw(),
            ffi::P3G_HASH_POSEIDON2_GL12,
            mats.as_ptr(),
            mats.len() as u64,
            log_blowup as u32,
            0,
            &mut raw,
        )
    };
    ffi::check(rc, "p3g_commit_traces");
    drop(guard);
    let count = traces.len();
    GpuTree::with_loader(
        raw,
        lmcs.ctx.clone(),
        Box::new(move |tree: &GpuTree<RowMajorMatrix<Goldilocks>>| {
            let _guard = tree.ctx.lock.lock().unwrap();
            (0..count)
                .map(|i| {
                    let (mut ptr, mut h, mut w) = (core::ptr::null(), 0u64, 0u64);
                    // SAFETY: i < the tree's matrix count; the device matrix has h * w words.
                    unsafe {
                        ffi::check(ffi::p3g_tree_matrix(tree.raw, i as u64, &mut ptr, &mut h, &mut w), "p3g_tree_matrix");
                        let mut host = vec![0u64; (h * w) as usize];
                        ffi::check(
                            ffi::p3g_download(tree.ctx.raw(), host.as_mut_ptr().cast(), ptr.cast(), 8 * h * w),
                            "p3g_download",
                        );
                        RowMajorMatrix::new(from_words(&host), w as usize)
                    }
                })
                .collect()
        }),
    )
}

/// The commitment of a tree, as the reference's `LmcsTree::root` returns it.
pub fn root<M: Matrix<Goldilocks>>(tree: &GpuTree<M>) -> Digest {
    LmcsTree::root(tree)
}
