//! The caller's `ProverChannel` (p3-miden-transcript/src/prover.rs:53-95) as the
//! library's `p3g_channel` callbacks.  The library calls them in exactly the order
//! the reference's `LmcsTree::prove_batch` and `open_with_channel` perform the
//! operations, so the caller's transcript records the reference's stream.

use core::ffi::{c_int, c_void};
use std::panic::{AssertUnwindSafe, catch_unwind};

use p3_field::PrimeField64;
use p3_goldilocks::Goldilocks;
use p3_miden_transcript::{Channel, ProverChannel};
use p3_symmetric::Hash;

use crate::dft::from_words;
use crate::ffi;

/// Commitment type of the Goldilocks Poseidon2 LMCS (`LmcsConfig::Commitment`, lmcs.rs:91).
pub type Digest = Hash<Goldilocks, Goldilocks, 4>;

/// A channel the callbacks can drive.
pub trait GpuChannel: ProverChannel<F = Goldilocks, Commitment = Digest> {}
impl<Ch: ProverChannel<F = Goldilocks, Commitment = Digest>> GpuChannel for Ch {}

fn digest(words: *const u64) -> Digest {
    // SAFETY: the library passes 4 canonical words for every commitment.
    let w = unsafe { core::slice::from_raw_parts(words, 4) };
    let f: [Goldilocks; 4] = core::array::from_fn(|i| Goldilocks::new(w[i]));
    Hash::from(f)
}

/// Runs `f` on the channel behind `user`; a panic becomes a nonzero status (the
/// library then aborts the call with P3G_E_TRANSCRIPT) instead of unwinding
/// through the C frames.
fn with<Ch: GpuChannel>(user: *mut c_void, f: impl FnOnce(&mut Ch)) -> c_int {
    // SAFETY: `user` is the `&mut Ch` that `bridge` stored, alive for the whole call.
    let ch = unsafe { &mut *(user as *mut Ch) };
    match catch_unwind(AssertUnwindSafe(|| f(ch))) {
        Ok(()) => 0,
        Err(_) => 1,
    }
}

extern "C" fn send_fields<Ch: GpuChannel>(u: *mut c_void, v: *const u64, n: u64) -> c_int {
    // SAFETY: the library passes n readable words.
    let vals = from_words(unsafe { core::slice::from_raw_parts(v, n as usize) });
    with::<Ch>(u, |ch| ch.send_field_slice(&vals))
}

extern "C" fn send_commitment<Ch: GpuChannel>(u: *mut c_void, c: *const u64) -> c_int {
    let d = digest(c);
    with::<Ch>(u, |ch| ch.send_commitment(d))
}

extern "C" fn hint_fields<Ch: GpuChannel>(u: *mut c_void, v: *const u64, n: u64) -> c_int {
    // SAFETY: the library passes n readable words.
    let vals = from_words(unsafe { core::slice::from_raw_parts(v, n as usize) });
    with::<Ch>(u, |ch| ch.hint_field_slice(&vals))
}

extern "C" fn hint_commitment<Ch: GpuChannel>(u: *mut c_void, c: *const u64) -> c_int {
    let d = digest(c);
    with::<Ch>(u, |ch| ch.hint_commitment(d))
}

extern "C" fn grind<Ch: GpuChannel>(u: *mut c_void, bits: u32, witness: *mut u64) -> c_int {
    with::<Ch>(u, |ch| {
        let w = ch.grind(bits as usize);
        // SAFETY: the library passes a writable word.
        unsafe { *witness = w.as_canonical_u64() };
    })
}

extern "C" fn sample<Ch: GpuChannel>(u: *mut c_void, out: *mut u64) -> c_int {
    with::<Ch>(u, |ch| {
        let v = Channel::sample(ch);
        // SAFETY: the library passes a writable word.
        unsafe { *out = v.as_canonical_u64() };
    })
}

extern "C" fn sample_bits<Ch: GpuChannel>(u: *mut c_void, bits: u32, out: *mut u64) -> c_int {
    with::<Ch>(u, |ch| {
        let v = Channel::sample_bits(ch, bits as usize);
        // SAFETY: the library passes a writable word.
        unsafe { *out = v as u64 };
    })
}

/// The callbacks over `ch`; the result must not outlive the borrow.
pub(crate) fn bridge<Ch: GpuChannel>(ch: &mut Ch) -> ffi::p3g_channel {
    ffi::p3g_channel {
        user: ch as *mut Ch as *mut c_void,
        send_fields: send_fields::<Ch>,
        send_commitment: send_commitment::<Ch>,
        hint_fields: hint_fields::<Ch>,
        hint_commitment: hint_commitment::<Ch>,
        grind: grind::<Ch>,
        sample: sample::<Ch>,
        sample_bits: sample_bits::<Ch>,
    }
}
