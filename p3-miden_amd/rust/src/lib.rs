//! Drop-in GPU backend for the lifted-STARK prover core of adr1anh/p3-miden over
//! Goldilocks, on one MI355X through `libp3gpu.so`:
//!
//! - [`GpuDft`] implements `TwoAdicSubgroupDft<Goldilocks>` (the `Dft` of a
//!   `StarkConfig`, p3-miden-lifted-stark/src/config.rs:26-42);
//! - [`GpuLmcs`] implements `Lmcs` with trees built and opened on the device (the
//!   config's `Lmcs`), wrapping the CPU `LmcsConfig` for the verifier side;
//! - [`open_with_channel`] is the PCS opening of p3-miden-lifted-fri/src/prover.rs:35-116
//!   with the same arguments, every transcript operation going to the caller's channel;
//! - [`commit_traces`] is commit.rs:175-228 in one call, the LDEs staying in HBM;
//! - [`GpuQuotient`] holds the quotient accumulator of `prove_multi` in HBM and
//!   evaluates an AIR's constraints on it ([`AirProgram`], constraints/mod.rs:63-179),
//!   reading the committed LDEs in place.
//!
//! ```ignore
//! let ctx = GpuContext::new(0);
//! let lmcs = GpuLmcs::new(LmcsConfig::new(sponge, compress), ctx.clone());
//! let config = GenericStarkConfig::new(pcs, lmcs, GpuDft::new(ctx), challenger);
//! ```
//!
//! Every proof equals the CPU prover's word for word: the library is tested against a
//! restatement of the reference prover.  Failures the reference reports by panicking
//! (unsorted or non-power-of-two heights, empty batches, out-of-range indices) panic
//! here with the library's message.

mod channel;
mod constraints;
mod context;
mod dft;
pub mod ffi;
mod lmcs;
mod open;

pub use channel::{Digest, GpuChannel};
pub use constraints::{AirProgram, AirValue, GpuQuotient};
pub use context::GpuContext;
pub use dft::GpuDft;
pub use lmcs::{GpuLmcs, GpuTree};
pub use open::{Challenge, GpuPcsParams, commit_traces, open_with_channel, root};
