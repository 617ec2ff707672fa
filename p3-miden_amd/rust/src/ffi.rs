//! Declarations of the C ABI in `include/p3gpu.h` that the shim calls, with the
//! same names, argument order and meaning.  Status codes: 0 is success; every
//! other value is one of the `P3G_E_*` codes and `p3g_last_error` holds the detail.
#![allow(non_camel_case_types, dead_code)]

use core::ffi::{c_char, c_int, c_void};

pub const P3G_OK: c_int = 0;
pub const P3G_E_INVAL: c_int = 1;
pub const P3G_E_NOT_POW2: c_int = 2;
pub const P3G_E_UNSORTED: c_int = 3;
pub const P3G_E_NOMEM: c_int = 4;
pub const P3G_E_DEVICE: c_int = 5;
pub const P3G_E_EMPTY: c_int = 6;
pub const P3G_E_TRANSCRIPT: c_int = 7;
pub const P3G_E_RANGE: c_int = 8;

pub const P3G_HASH_POSEIDON2_GL12: u32 = 0;
pub const P3G_HASH_KECCAK_GL: u32 = 1;

pub const P3G_DEVICE_PTRS: u32 = 1;
pub const P3G_BITREV_ROWS: u32 = 2;

#[repr(C)]
pub struct p3g_ctx {
    _opaque: [u8; 0],
}

#[repr(C)]
pub struct p3g_tree {
    _opaque: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct p3g_matrix {
    pub data: *const u64,
    pub height: u64,
    pub width: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct p3g_pcs_params {
    pub log_blowup: u32,
    pub log_folding_arity: u32,
    pub log_final_degree: u32,
    pub folding_pow_bits: u32,
    pub deep_pow_bits: u32,
    pub num_queries: u32,
    pub query_pow_bits: u32,
}

/// One op of a constraint program (`p3g_air_op` in the header).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq, Eq)]
pub struct p3g_air_op {
    pub op: u32,
    pub a: u32,
    pub b: u32,
    pub reserved: u32,
    pub imm: u64,
}

pub const P3G_AIR_CONST: u32 = 0;
pub const P3G_AIR_MAIN: u32 = 1;
pub const P3G_AIR_AUX: u32 = 2;
pub const P3G_AIR_PERIODIC: u32 = 3;
pub const P3G_AIR_PUBLIC: u32 = 4;
pub const P3G_AIR_EXT_INPUT: u32 = 5;
pub const P3G_AIR_IS_FIRST_ROW: u32 = 6;
pub const P3G_AIR_IS_LAST_ROW: u32 = 7;
pub const P3G_AIR_IS_TRANSITION: u32 = 8;
pub const P3G_AIR_ADD: u32 = 9;
pub const P3G_AIR_SUB: u32 = 10;
pub const P3G_AIR_MUL: u32 = 11;
pub const P3G_AIR_NEG: u32 = 12;
pub const P3G_AIR_ASSERT_ZERO: u32 = 13;
pub const P3G_AIR_ASSERT_ZERO_EXT: u32 = 14;

/// The caller's ProverChannel as callbacks (`p3g_channel` in the header).
#[repr(C)]
pub struct p3g_channel {
    pub user: *mut c_void,
    pub send_fields: extern "C" fn(*mut c_void, *const u64, u64) -> c_int,
    pub send_commitment: extern "C" fn(*mut c_void, *const u64) -> c_int,
    pub hint_fields: extern "C" fn(*mut c_void, *const u64, u64) -> c_int,
    pub hint_commitment: extern "C" fn(*mut c_void, *const u64) -> c_int,
    pub grind: extern "C" fn(*mut c_void, u32, *mut u64) -> c_int,
    pub sample: extern "C" fn(*mut c_void, *mut u64) -> c_int,
    pub sample_bits: extern "C" fn(*mut c_void, u32, *mut u64) -> c_int,
}

unsafe extern "C" {
    pub fn p3g_status_string(status: c_int) -> *const c_char;
    pub fn p3g_last_error() -> *const c_char;

    pub fn p3g_ctx_create(device: c_int, out: *mut *mut p3g_ctx) -> c_int;
    pub fn p3g_ctx_destroy(ctx: *mut p3g_ctx);
    pub fn p3g_ctx_synchronize(ctx: *mut p3g_ctx) -> c_int;

    pub fn p3g_coset_lde_batch(
        ctx: *mut p3g_ctx,
        trace: *const u64,
        n: u64,
        w: u64,
        added_bits: u32,
        shift: u64,
        flags: u32,
        out: *mut u64,
    ) -> c_int;

    pub fn p3g_lmcs_alignment(hash_kind: u32) -> u32;
    pub fn p3g_lmcs_build(
        ctx: *mut p3g_ctx,
        hash_kind: u32,
        mats: *const p3g_matrix,
        k: u64,
        alignment: u32,
        flags: u32,
        out: *mut *mut p3g_tree,
    ) -> c_int;
    pub fn p3g_commit_traces(
        ctx: *mut p3g_ctx,
        hash_kind: u32,
        traces: *const p3g_matrix,
        k: u64,
        log_blowup: u32,
        flags: u32,
        out: *mut *mut p3g_tree,
    ) -> c_int;

    pub fn p3g_tree_root(t: *const p3g_tree, out: *mut u64) -> c_int;
    pub fn p3g_tree_height(t: *const p3g_tree) -> u64;
    pub fn p3g_tree_alignment(t: *const p3g_tree) -> u32;
    pub fn p3g_tree_num_matrices(t: *const p3g_tree) -> u64;
    pub fn p3g_tree_widths(t: *const p3g_tree, out: *mut u64) -> c_int;
    pub fn p3g_tree_rows(t: *mut p3g_tree, index: u64, out: *mut u64, cap: u64, len: *mut u64) -> c_int;
    pub fn p3g_tree_prove_batch_cb(t: *mut p3g_tree, indices: *const u64, n: u64, ch: *const p3g_channel) -> c_int;
    pub fn p3g_tree_matrix(t: *mut p3g_tree, i: u64, dptr: *mut *const u64, height: *mut u64, width: *mut u64)
    -> c_int;
    pub fn p3g_tree_free(t: *mut p3g_tree);
    pub fn p3g_download(ctx: *mut p3g_ctx, dst: *mut c_void, src: *const c_void, bytes: u64) -> c_int;
    pub fn p3g_alloc(ctx: *mut p3g_ctx, bytes: u64, dptr: *mut *mut c_void) -> c_int;
    pub fn p3g_free(ctx: *mut p3g_ctx, dptr: *mut c_void) -> c_int;

    pub fn p3g_quotient_extend_scale(ctx: *mut p3g_ctx, acc: *mut u64, m: u64, target: u64, beta: *const u64)
    -> c_int;
    pub fn p3g_quotient_divide_vanishing(ctx: *mut p3g_ctx, num: *mut u64, n: u64, d: u64, shift: u64) -> c_int;
    pub fn p3g_air_compile(ops: *const p3g_air_op, nops: u64, code_bytes: *mut u64) -> c_int;
    pub fn p3g_eval_constraints(
        ctx: *mut p3g_ctx,
        ops: *const p3g_air_op,
        nops: u64,
        main: *const p3g_matrix,
        aux: *const p3g_matrix,
        periodic: *const u64,
        periodic_rows: u64,
        periodic_cols: u64,
        publics: *const u64,
        num_publics: u64,
        ext_inputs: *const u64,
        num_ext_inputs: u64,
        n: u64,
        log_d: u32,
        shift: u64,
        alpha: *const u64,
        accumulator: *mut u64,
    ) -> c_int;

    pub fn p3g_open_with_callbacks(
        ctx: *mut p3g_ctx,
        params: *const p3g_pcs_params,
        log_lde_height: u32,
        eval_points: *const u64,
        n_points: u32,
        trees: *const *mut p3g_tree,
        ntrees: u64,
        ch: *const p3g_channel,
    ) -> c_int;
}

/// The library's message for a failed call.
pub fn last_error(status: c_int) -> String {
    // SAFETY: both functions return pointers to NUL-terminated strings owned by the library
    // (a static table and a thread-local buffer) that stay valid until the next call.
    unsafe {
        let what = core::ffi::CStr::from_ptr(p3g_status_string(status)).to_string_lossy();
        let detail = core::ffi::CStr::from_ptr(p3g_last_error()).to_string_lossy();
        format!("{what} ({status}): {detail}")
    }
}

/// Panics with the library's message unless `status` is success: the reference prover
/// panics on the same precondition violations (`lifted_tree.rs:174-177,262`,
/// `lifted-fri/src/prover.rs:53-60`).
#[track_caller]
pub fn check(status: c_int, call: &str) {
    if status != P3G_OK {
        panic!("{call}: {}", last_error(status));
    }
}
