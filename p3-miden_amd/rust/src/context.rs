//! One library context per device, shared by the DFT, the LMCS and the opening.

use std::sync::{Arc, Mutex, OnceLock};

use crate::ffi;

/// Owns a `p3g_ctx` (its HIP stream, device block cache and tables).  The library
/// serialises a context's work on its stream; the mutex keeps calls from several
/// Rust threads from interleaving their host-side steps on it.
pub struct GpuContext {
    raw: *mut ffi::p3g_ctx,
    pub(crate) lock: Mutex<()>,
}

// SAFETY: the context is only used under `lock`, and the library allows a context to
// be driven from any thread as long as its calls do not overlap.
unsafe impl Send for GpuContext {}
unsafe impl Sync for GpuContext {}

impl GpuContext {
    pub fn new(device: i32) -> Arc<Self> {
        let mut raw = core::ptr::null_mut();
        // SAFETY: `raw` is a valid out-pointer; on success the library hands over ownership.
        ffi::check(unsafe { ffi::p3g_ctx_create(device, &mut raw) }, "p3g_ctx_create");
        Arc::new(Self {
            raw,
            lock: Mutex::new(()),
        })
    }

    /// The context of device 0, created on first use (what `Default` impls share).
    pub fn default_device() -> Arc<Self> {
        static CTX: OnceLock<Arc<GpuContext>> = OnceLock::new();
        CTX.get_or_init(|| GpuContext::new(0)).clone()
    }

    pub(crate) fn raw(&self) -> *mut ffi::p3g_ctx {
        self.raw
    }
}

impl Drop for GpuContext {
    fn drop(&mut self) {
        // SAFETY: every tree that refers to this context holds an Arc to it, so none is alive.
        unsafe { ffi::p3g_ctx_destroy(self.raw) }
    }
}
