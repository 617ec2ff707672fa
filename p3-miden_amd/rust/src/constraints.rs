//! Constraint evaluation on the quotient domain (`evaluate_constraints_into`,
//! p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179) and the quotient
//! accumulator around it (`cyclic_extend_and_scale`, `divide_by_vanishing_in_place`,
//! quotient.rs:40-101), with the accumulator in HBM and the traces read in place from
//! the trees of [`crate::commit_traces`].
//!
//! The AIR crosses the boundary as data: an [`AirProgram`], the straight-line record of
//! its `eval` (one op per value, children first).  A symbolic AIR builder's constraint
//! lists map onto it node by node; the library compiles each program once into its own
//! gfx950 kernel.

use core::ffi::c_void;
use std::sync::Arc;

use p3_field::{BasedVectorSpace, PrimeField64};
use p3_goldilocks::Goldilocks;
use p3_matrix::Matrix;

use crate::context::GpuContext;
use crate::dft::to_words;
use crate::ffi;
use crate::lmcs::GpuTree;
use crate::open::Challenge;

/// A value of an [`AirProgram`]: the index of the op that computes it.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct AirValue(u32);

/// An AIR's constraints as a straight-line program (`p3g_air_op`, include/p3gpu.h), in
/// the AirBuilder vocabulary: the main and permutation (aux) windows on the current or
/// next row, periodic values, public values, extension inputs (the permutation
/// randomness, then the permutation values), the three selectors, ring operations, and
/// the assertions in the order the AIR emits them (the fold Σ α^(K−1−k)·C_k follows it).
#[derive(Clone, Debug, Default)]
pub struct AirProgram {
    ops: Vec<ffi::p3g_air_op>,
}

impl AirProgram {
    pub fn new() -> Self {
        Self::default()
    }

    fn push(&mut self, op: u32, a: u32, b: u32, imm: u64) -> AirValue {
        self.ops.push(ffi::p3g_air_op { op, a, b, reserved: 0, imm });
        AirValue((self.ops.len() - 1) as u32)
    }

    pub fn constant(&mut self, c: Goldilocks) -> AirValue {
        self.push(ffi::P3G_AIR_CONST, 0, 0, c.as_canonical_u64())
    }
    /// `main().current_slice()[col]`, or `next_slice()[col]`.
    pub fn main(&mut self, col: usize, next: bool) -> AirValue {
        self.push(ffi::P3G_AIR_MAIN, col as u32, next as u32, 0)
    }
    /// Extension column `col` of the permutation window.
    pub fn aux(&mut self, col: usize, next: bool) -> AirValue {
        self.push(ffi::P3G_AIR_AUX, col as u32, next as u32, 0)
    }
    pub fn periodic(&mut self, col: usize) -> AirValue {
        self.push(ffi::P3G_AIR_PERIODIC, col as u32, 0, 0)
    }
    pub fn public(&mut self, i: usize) -> AirValue {
        self.push(ffi::P3G_AIR_PUBLIC, i as u32, 0, 0)
    }
    /// `permutation_randomness()[i]`; the permutation values follow the randomness.
    pub fn ext_input(&mut self, i: usize) -> AirValue {
        self.push(ffi::P3G_AIR_EXT_INPUT, i as u32, 0, 0)
    }
    pub fn is_first_row(&mut self) -> AirValue {
        self.push(ffi::P3G_AIR_IS_FIRST_ROW, 0, 0, 0)
    }
    pub fn is_last_row(&mut self) -> AirValue {
        self.push(ffi::P3G_AIR_IS_LAST_ROW, 0, 0, 0)
    }
    pub fn is_transition(&mut self) -> AirValue {
        self.push(ffi::P3G_AIR_IS_TRANSITION, 0, 0, 0)
    }
    pub fn add(&mut self, x: AirValue, y: AirValue) -> AirValue {
        self.push(ffi::P3G_AIR_ADD, x.0, y.0, 0)
    }
    pub fn sub(&mut self, x: AirValue, y: AirValue) -> AirValue {
        self.push(ffi::P3G_AIR_SUB, x.0, y.0, 0)
    }
    pub fn mul(&mut self, x: AirValue, y: AirValue) -> AirValue {
        self.push(ffi::P3G_AIR_MUL, x.0, y.0, 0)
    }
    pub fn neg(&mut self, x: AirValue) -> AirValue {
        self.push(ffi::P3G_AIR_NEG, x.0, 0, 0)
    }
    pub fn assert_zero(&mut self, x: AirValue) {
        self.push(ffi::P3G_AIR_ASSERT_ZERO, x.0, 0, 0);
    }
    pub fn assert_zero_ext(&mut self, x: AirValue) {
        self.push(ffi::P3G_AIR_ASSERT_ZERO_EXT, x.0, 0, 0);
    }

    pub fn ops(&self) -> &[ffi::p3g_air_op] {
        &self.ops
    }

    /// Checks the program and compiles its kernel ahead of the first evaluation (no
    /// device needed); returns the code object's size.  Panics with the library's
    /// message on a malformed program.
    pub fn compile(&self) -> usize {
        let mut size = 0u64;
        // SAFETY: `ops` is a live slice of `p3g_air_op` and `size` a valid output.
        ffi::check(
            unsafe { ffi::p3g_air_compile(self.ops.as_ptr(), self.ops.len() as u64, &mut size) },
            "p3g_air_compile",
        );
        size as usize
    }
}

fn ef_words(values: &[Challenge]) -> Vec<u64> {
    values
        .iter()
        .flat_map(|e| e.as_basis_coefficients_slice().iter().map(|c| c.as_canonical_u64()))
        .collect()
}

/// The quotient accumulator of `prove_multi` (prover/mod.rs:358-432) in HBM: up to
/// `capacity` extension elements, of which the first `len` are live.
pub struct GpuQuotient {
    ctx: Arc<GpuContext>,
    ptr: *mut u64,
    capacity: usize,
    len: usize,
}

// SAFETY: the buffer is only used under the context's lock.
unsafe impl Send for GpuQuotient {}
unsafe impl Sync for GpuQuotient {}

impl Drop for GpuQuotient {
    fn drop(&mut self) {
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: `ptr` came from p3g_alloc on this context and is freed once.
        unsafe { ffi::p3g_free(self.ctx.raw(), self.ptr as *mut c_void) };
    }
}

/// The device view of matrix `i` of a tree: its bit-reversed LDE rows.
fn tree_matrix<M>(tree: &GpuTree<M>, i: usize) -> ffi::p3g_matrix {
    let (mut data, mut height, mut width) = (core::ptr::null(), 0u64, 0u64);
    // SAFETY: the tree handle is live; the outputs are valid.
    ffi::check(
        unsafe { ffi::p3g_tree_matrix(tree.raw, i as u64, &mut data, &mut height, &mut width) },
        "p3g_tree_matrix",
    );
    ffi::p3g_matrix { data, height, width }
}

impl GpuQuotient {
    /// An empty accumulator with room for `capacity` (the largest N·D) extension elements.
    pub fn new(ctx: Arc<GpuContext>, capacity: usize) -> Self {
        let mut ptr = core::ptr::null_mut::<c_void>();
        {
            let _guard = ctx.lock.lock().unwrap();
            // SAFETY: the context is live; `ptr` is a valid output.
            ffi::check(unsafe { ffi::p3g_alloc(ctx.raw(), 16 * capacity as u64, &mut ptr) }, "p3g_alloc");
        }
        Self { ctx, ptr: ptr as *mut u64, capacity, len: 0 }
    }

    pub fn len(&self) -> usize {
        self.len
    }

    /// The device buffer (`len` extension elements as (c0, c1) words), for
    /// `p3g_commit_quotient` with `P3G_DEVICE_PTRS`.
    pub fn as_device_ptr(&self) -> *const u64 {
        self.ptr
    }

    /// `cyclic_extend_and_scale` (quotient.rs:40-55): an empty accumulator becomes
    /// `target` zeros; otherwise it is scaled by `beta` and repeated up to `target`.
    pub fn extend_scale(&mut self, target: usize, beta: Challenge) {
        assert!(target <= self.capacity, "quotient accumulator capacity {} < {target}", self.capacity);
        let b = ef_words(&[beta]);
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: the buffer holds `capacity` >= `target` elements; `b` has 2 words.
        ffi::check(
            unsafe {
                ffi::p3g_quotient_extend_scale(self.ctx.raw(), self.ptr, self.len as u64, target as u64, b.as_ptr())
            },
            "p3g_quotient_extend_scale",
        );
        self.len = target;
    }

    /// `evaluate_constraints_into` for one instance: matrix `main.1` of tree `main.0` (and
    /// `aux.1` of `aux.0`, whose extension columns are committed as (c0, c1) base pairs)
    /// read in place; `periodic` the PeriodicLde values on gJ (natural order, `width`
    /// columns); `ext_inputs` the permutation randomness, then the permutation values.
    #[allow(clippy::too_many_arguments)]
    pub fn eval_constraints<M: Matrix<Goldilocks>, A: Matrix<Goldilocks>>(
        &mut self,
        program: &AirProgram,
        main: (&GpuTree<M>, usize),
        aux: Option<(&GpuTree<A>, usize)>,
        periodic: Option<(&[Goldilocks], usize)>,
        publics: &[Goldilocks],
        ext_inputs: &[Challenge],
        trace_height: usize,
        log_quotient_degree: usize,
        shift: Goldilocks,
        alpha: Challenge,
    ) {
        assert_eq!(self.len, trace_height << log_quotient_degree, "accumulator not extended to N*D");
        let m = tree_matrix(main.0, main.1);
        let x = aux.map(|(t, i)| tree_matrix(t, i));
        let per = periodic.map(|(v, w)| (to_words(v), w));
        let (per_ptr, per_rows, per_cols) = match &per {
            Some((v, w)) => (v.as_ptr(), (v.len() / w) as u64, *w as u64),
            None => (core::ptr::null(), 0, 0),
        };
        let pubs = to_words(publics);
        let ext = ef_words(ext_inputs);
        let al = ef_words(&[alpha]);
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: every pointer is live for the call and sized as passed; the matrices are
        // the trees' device buffers and the accumulator holds `len` elements.
        ffi::check(
            unsafe {
                ffi::p3g_eval_constraints(
                    self.ctx.raw(),
                    program.ops.as_ptr(),
                    program.ops.len() as u64,
                    &m,
                    x.as_ref().map_or(core::ptr::null(), |x| x as *const _),
                    per_ptr,
                    per_rows,
                    per_cols,
                    pubs.as_ptr(),
                    pubs.len() as u64,
                    ext.as_ptr(),
                    ext_inputs.len() as u64,
                    trace_height as u64,
                    log_quotient_degree as u32,
                    shift.as_canonical_u64(),
                    al.as_ptr(),
                    self.ptr,
                )
            },
            "p3g_eval_constraints",
        );
    }

    /// `divide_by_vanishing_in_place` (quotient.rs:72-101) on the N·D values of the
    /// coset `shift`·J.
    pub fn divide_by_vanishing(&mut self, trace_height: usize, quotient_degree: usize, shift: Goldilocks) {
        assert_eq!(self.len, trace_height * quotient_degree);
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: the buffer holds `len` = n d elements.
        ffi::check(
            unsafe {
                ffi::p3g_quotient_divide_vanishing(
                    self.ctx.raw(),
                    self.ptr,
                    trace_height as u64,
                    quotient_degree as u64,
                    shift.as_canonical_u64(),
                )
            },
            "p3g_quotient_divide_vanishing",
        );
    }

    /// The live values, copied to the host.
    pub fn to_host(&self) -> Vec<Challenge> {
        let mut words = vec![0u64; 2 * self.len];
        let _guard = self.ctx.lock.lock().unwrap();
        // SAFETY: `words` has room for the `len` elements read.
        ffi::check(
            unsafe {
                ffi::p3g_download(
                    self.ctx.raw(),
                    words.as_mut_ptr() as *mut c_void,
                    self.ptr as *const c_void,
                    16 * self.len as u64,
                )
            },
            "p3g_download",
        );
        words
            .chunks_exact(2)
            .map(|w| Challenge::from_basis_coefficients_fn(|i| Goldilocks::new(w[i])))
            .collect()
    }
}
