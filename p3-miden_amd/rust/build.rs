// Links libp3gpu.so, built by `make -C p3-miden_amd` (hipcc, gfx950).  P3GPU_DIR
// names the directory that holds it; the runtime loader must find it too
// (LD_LIBRARY_PATH or an rpath).
fn main() {
    let dir = std::env::var("P3GPU_DIR").unwrap_or_else(|_| "../p3-miden_amd".to_string());
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=p3gpu");
    println!("cargo:rerun-if-env-changed=P3GPU_DIR");
}
