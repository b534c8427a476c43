//! Links libzfp_hip.so: ZFP_HIP_DIR names the directory holding it (the
//! `zfp-rs_amd/` folder after `make -C zfp-rs_amd`).
fn main() {
    let dir = std::env::var("ZFP_HIP_DIR").unwrap_or_else(|_| "..".to_string());
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=zfp_hip");
    println!("cargo:rerun-if-env-changed=ZFP_HIP_DIR");
}
